// Path-guiding network for gfx950: a 48 -> 64 -> 64 -> 64 -> 64 -> 40 ReLU MLP in BF16 with FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), whose 40 outputs parameterise a
// mixture of 8 von Mises-Fisher lobes (3 direction + concentration + weight logit per lobe).
// No counterpart exists in the reference renderer; the network only changes which directions are
// sampled, and one-sample MIS with the BSDF/phase density keeps the estimator unbiased.
//
// Matrix layout. Activations are kept "neuron-major": an MFMA output tile D[16 neurons][16
// samples] leaves lane l with D[4*(l/16) + j][l%16], j = 0..3. Two such tiles (neurons n0..n0+15
// and n0+16..n0+31) are exactly the B operand of the next layer's K = 32 step when the K index
// is permuted as k(g, j) = (j < 4 ? 4g + j : 16 + 4g + j - 4), g = l/16; the weight fragments (A
// operand, lane l holds A[l%16][k(g, 0..7)]) are stored pre-permuted, so hidden activations never
// leave registers between layers.
#pragma once

#include <cstdint>

namespace trg {

constexpr int kFeat = 48;      // encoded input features
constexpr int kHid = 64;       // hidden width
constexpr int kLobes = 8;
constexpr int kOut = 5 * kLobes;  // 40 network outputs
constexpr int kLayers = 5;
constexpr int kRaw = 12;       // raw vertex inputs: pos(3, scene-normalised), wo(3), n(3), t(1), medium(1), pad
// per layer: inputs, outputs, padded outputs (multiple of 16) and fragment counts
constexpr int kIn[kLayers] = {kFeat, kHid, kHid, kHid, kHid};
constexpr int kOutL[kLayers] = {kHid, kHid, kHid, kHid, kOut};
constexpr int kOutPad[kLayers] = {64, 64, 64, 64, 48};
// master FP32 parameters: per layer W[in][out] (in-major) then b[out]
constexpr int param_offset(int l)
{
    int o = 0;
    for (int i = 0; i < l; ++i) o += kIn[i] * kOutL[i] + kOutL[i];
    return o;
}
constexpr int kParams = param_offset(kLayers);  // 18,216
constexpr int kWeights = kFeat * kHid + 3 * kHid * kHid + kHid * kOut;  // 17,920
// forward fragments (A = W^T, 16 outputs x 32 permuted inputs, 64 lanes x 8 bf16 = 1 KiB each)
constexpr int fwd_frag_offset(int l)
{
    int o = 0;
    for (int i = 0; i < l; ++i) o += (kOutPad[i] / 16) * 2;
    return o;
}
constexpr int kFwdFrags = fwd_frag_offset(kLayers);  // 38
// backward fragments (A = W, 16 inputs x 32 permuted outputs) for layers 1..4, outputs padded to 64
constexpr int kBwdFragsPerLayer = 4 * 2;
constexpr int kBwdFrags = 4 * kBwdFragsPerLayer;  // 32
constexpr int kBiasPad = kLayers * 64;
// FLOP per evaluated sample (2 x multiply-accumulates of the unpadded network)
constexpr double kFlopForward = 2.0 * kWeights;  // 35,840

// permuted K index of lane group g, element j (see the header comment)
constexpr int kperm(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }

// round-to-nearest-even FP32 -> BF16 bits (inputs are finite)
inline uint16_t bf16_bits_host(float x)
{
    union {
        float f;
        uint32_t u;
    } v{x};
    uint32_t u = v.u;
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

}  // namespace trg

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

namespace trg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// float -> bf16 bits, rounded to nearest even by the hardware conversion (v_cvt_pk_bf16_f32, two
// values per instruction): the same bits as bf16_bits_host for every finite input
__device__ __forceinline__ uint32_t bf16_bits(float x) { return __builtin_bit_cast(uint16_t, __bf16(x)); }
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi)
{
    const bf16x2 v = {__bf16(lo), __bf16(hi)};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ __forceinline__ float bf16_to_f32(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

__device__ __forceinline__ f32x4 mfma32(uint4 a, uint4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// Encoded features of one vertex (48 of them; see the host restatement in the oracle tests):
//   pos (3, normalised to the scene box) and sin/cos(2^f * pi * pos) for f = 0..3      27
//   wo (3) and the shading normal (3, zero at a medium vertex)                          6
//   t/total_time and sin/cos(2^f * pi * t) for f = 0..4                                11
//   medium flag                                                                         1
//   zero padding                                                                        3
// sin(pi 2^q x), cos(pi 2^q x) for q = 0 .. Q-1 into f[2q], f[2q + 1]: the hardware sine and
// cosine (v_sin_f32 / v_cos_f32 take their argument in turns; the fraction keeps it inside their
// valid range), then the double-angle identities sin 2a = 2 sin a cos a, cos 2a = 1 - 2 sin^2 a
// (each doubling at most doubles the absolute error: the top octave stays within ~1e-5, far below
// the bf16 rounding the network's inputs get)
template <int Q>
__device__ __forceinline__ void octaves(float x, float* f)
{
    const float turns = __builtin_amdgcn_fractf(0.5f * x);
    float sn = __builtin_amdgcn_sinf(turns), cs = __builtin_amdgcn_cosf(turns);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        f[2 * q] = sn;
        f[2 * q + 1] = cs;
        const float s2 = 2.0f * sn * cs, c2 = fmaf(-2.0f * sn, sn, 1.0f);
        sn = s2;
        cs = c2;
    }
}

__device__ __forceinline__ void encode_features(const float* raw, float* f)
{
    int k = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[k++] = raw[c];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        octaves<4>(raw[c], f + k);
        k += 8;
    }
#pragma unroll
    for (int c = 3; c < 9; ++c) f[k++] = raw[c];
    f[k++] = raw[9];
    octaves<5>(raw[9], f + k);
    k += 10;
    f[k++] = raw[10];
    while (k < kFeat) f[k++] = 0.0f;
}

// Per-wavefront LDS block of the network: 64 rows of kRow words. Row s first holds sample s's
// 64 bf16 features (32 words, 16 zero columns), later its 40 FP32 outputs. The row stride is odd
// (41 words) so that the 64 lanes reading word k of their own rows hit 64 different banks (a
// stride of 40 words put them on 8 banks: 8-way conflicts on every read of the outputs).
constexpr int kRow = kOut + 1;

__device__ __forceinline__ void store_feature_row(uint32_t* row32, const float* f)
{
#pragma unroll
    for (int i = 0; i < kFeat / 2; ++i) row32[i] = pack_bf16(f[2 * i], f[2 * i + 1]);
#pragma unroll
    for (int i = kFeat / 2; i < 32; ++i) row32[i] = 0u;
}

// One wavefront evaluates the network for 64 samples held in `buf` (64 x kRow words, see above),
// NST tiles of 16 samples per pass (more tiles: fewer weight-fragment reads, more registers).
// `frag` holds the forward fragments (64 lanes x uint4 each; LDS or global memory), `bias` the
// padded biases (64 per layer; LDS or global memory). The 40 outputs of sample s replace its features in row s; a row
// is overwritten only after its pass has read it. Every lane of the wavefront must call this.
// If `acts` is not null (training), the neuron-major bf16 inputs of layers 1..4 are also stored:
// acts[(l - 1) * 64 * 64 + n * 64 + s]. Only the sample tiles [st_begin, st_end) are evaluated
// (the training kernel splits a tile's four sample tiles over four wavefronts).
template <int NST>
__device__ __forceinline__ void mlp_forward_wave(float* buf, const uint4* frag, const float* bias, uint16_t* acts,
                                                 int st_begin = 0, int st_end = 4)
{
    const int lane = int(__lane_id());
    const int g = lane >> 4, c = lane & 15;
    const uint32_t* rows = reinterpret_cast<const uint32_t*>(buf);
#pragma unroll 1
    for (int st0 = st_begin; st0 < st_end; st0 += NST) {
        // B operands of layer 0 straight from the feature rows: sample (st*16 + c), inputs k(g, j)
        uint4 B[NST][2];
#pragma unroll
        for (int t = 0; t < NST; ++t)
#pragma unroll
            for (int kp = 0; kp < 2; ++kp) {
                // 32-bit loads: rows of an odd word count are only 4-byte aligned
                const uint32_t* r = rows + ((st0 + t) * 16 + c) * kRow + kp * 16;
                B[t][kp] = make_uint4(r[2 * g], r[2 * g + 1], r[8 + 2 * g], r[9 + 2 * g]);
            }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int l = 0; l < kLayers; ++l) {
            const int ntiles = kOutPad[l] / 16;
            f32x4 acc[4][NST];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                if (nt >= ntiles) break;
                float4 b4 = *reinterpret_cast<const float4*>(bias + l * 64 + nt * 16 + 4 * g);
#pragma unroll
                for (int t = 0; t < NST; ++t) acc[nt][t] = f32x4{b4.x, b4.y, b4.z, b4.w};
            }
#pragma unroll
            for (int kp = 0; kp < 2; ++kp)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    if (nt >= ntiles) break;
                    uint4 a = frag[(fwd_frag_offset(l) + nt * 2 + kp) * 64 + lane];
#pragma unroll
                    for (int t = 0; t < NST; ++t) acc[nt][t] = mfma32(a, B[t][kp], acc[nt][t]);
                }
            if (l < kLayers - 1) {
                // ReLU, round to bf16, and re-pair the tiles as the next layer's B operands
#pragma unroll
                for (int t = 0; t < NST; ++t)
#pragma unroll
                    for (int kp = 0; kp < 2; ++kp) {
                        f32x4 x = acc[2 * kp][t], y = acc[2 * kp + 1][t];
                        uint4 v;
                        v.x = pack_bf16(fmaxf(x[0], 0.f), fmaxf(x[1], 0.f));
                        v.y = pack_bf16(fmaxf(x[2], 0.f), fmaxf(x[3], 0.f));
                        v.z = pack_bf16(fmaxf(y[0], 0.f), fmaxf(y[1], 0.f));
                        v.w = pack_bf16(fmaxf(y[2], 0.f), fmaxf(y[3], 0.f));
                        B[t][kp] = v;
                        if (acts) {
                            uint16_t* A = acts + l * 64 * 64;
                            const int s = (st0 + t) * 16 + c;
                            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                int n = kp * 32 + kperm(g, j);
                                A[n * 64 + s] = uint16_t((w4[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu);
                            }
                        }
                    }
            } else {
#pragma unroll
                for (int nt = 0; nt < 3; ++nt)
#pragma unroll
                    for (int t = 0; t < NST; ++t)
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            int o = nt * 16 + 4 * g + j;
                            if (o < kOut) buf[((st0 + t) * 16 + c) * kRow + o] = acc[nt][t][j];
                        }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- von Mises-Fisher mixture ---------------------------------------------------------------
struct Lobe {
    float mx, my, mz, kappa, w;
};

__device__ __forceinline__ float softplus(float x) { return x > 15.f ? x : log1pf(__expf(x)); }
// the same from the hardware exp / log (shading: kappa is a parameter of the proposal that is both
// sampled and evaluated, so its last digits need not match the training kernel's)
__device__ __forceinline__ float softplus_fast(float x) { return x > 15.f ? x : __logf(1.0f + __expf(x)); }

constexpr float kKappaMin = 1e-2f;
constexpr float kKappaMax = 2e3f;

// The mixture of one sample, decoded once from its 40 raw outputs (kept in LDS): per lobe the unit
// mean direction, kappa, the vMF density's normaliser kappa / (2 pi (1 - e^{-2 kappa})) (the
// density is norm * exp(kappa (mu.w - 1))) and the softmax weight, in registers, so that picking a
// lobe, sampling it and evaluating the mixture density decode nothing again. The decode uses the
// hardware exp, log and reciprocal (about 1 ulp): with kappa >= 0.01, 1 - e^{-2 kappa} >= 0.0198
// keeps the normaliser within ~1e-5 relative without expm1, and the density that weights a sample
// is the one it was drawn from up to that.
struct Mixture {
    float mx[kLobes], my[kLobes], mz[kLobes], kappa[kLobes], norm[kLobes], w[kLobes];
};
__device__ __forceinline__ void mixture_decode(const float* o, Mixture& M)
{
    float lmax = -3.4e38f;
#pragma unroll
    for (int i = 0; i < kLobes; ++i) lmax = fmaxf(lmax, o[5 * i + 4]);
    float sum = 0.f;
#pragma unroll
    for (int i = 0; i < kLobes; ++i) {
        M.w[i] = __expf(o[5 * i + 4] - lmax);
        sum += M.w[i];
    }
    const float inv_sum = __builtin_amdgcn_rcpf(sum);
#pragma unroll
    for (int i = 0; i < kLobes; ++i) {
        M.w[i] *= inv_sum;
        const float x = o[5 * i], y = o[5 * i + 1], z = o[5 * i + 2];
        const float r2 = x * x + y * y + z * z;
        if (r2 > 1e-38f) {
            const float ir = rsqrtf(r2);
            M.mx[i] = x * ir, M.my[i] = y * ir, M.mz[i] = z * ir;
        } else {
            M.mx[i] = 0.f, M.my[i] = 0.f, M.mz[i] = 1.f;
        }
        const float k = fminf(softplus_fast(o[5 * i + 3]) + kKappaMin, kKappaMax);
        M.kappa[i] = k;
        M.norm[i] = k * __builtin_amdgcn_rcpf(2.0f * 3.14159265358979f * (1.0f - __expf(-2.0f * k)));
    }
}
// lobe i of the mixture (i uniform or not: selected without indexing the register arrays)
__device__ __forceinline__ Lobe lobe_of(const Mixture& M, int i)
{
    Lobe L{M.mx[0], M.my[0], M.mz[0], M.kappa[0], M.w[0]};
#pragma unroll
    for (int j = 1; j < kLobes; ++j)
        if (i == j) L = Lobe{M.mx[j], M.my[j], M.mz[j], M.kappa[j], M.w[j]};
    return L;
}
__device__ __forceinline__ float mixture_pdf(const Mixture& M, float wx, float wy, float wz)
{
    float p = 0.f;
#pragma unroll
    for (int i = 0; i < kLobes; ++i) {
        const float c = M.mx[i] * wx + M.my[i] * wy + M.mz[i] * wz;
        p += M.w[i] * M.norm[i] * __expf(M.kappa[i] * (c - 1.0f));
    }
    return p;
}
// lobe index for a uniform number u (inverse CDF of the weights)
__device__ __forceinline__ int pick_lobe(const Mixture& M, float u)
{
    float acc = 0.f;
    int k = kLobes - 1;
    bool found = false;
#pragma unroll
    for (int i = 0; i < kLobes - 1; ++i) {
        acc += M.w[i];
        if (!found && u < acc) {
            k = i;
            found = true;
        }
    }
    return k;
}

// Sample a vMF lobe (W. Jakob, "Numerically stable sampling of the von Mises-Fisher distribution
// on S^2", 2012): cos = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa, azimuth 2 pi u2 (hardware
// log, reciprocal, and sine / cosine, whose argument is in turns: u2 itself).
__device__ __forceinline__ void vmf_sample(const Lobe& l, float u1, float u2, float* wx, float* wy, float* wz)
{
    float ct = 1.0f + __logf(u1 + (1.0f - u1) * __expf(-2.0f * l.kappa)) * __builtin_amdgcn_rcpf(l.kappa);
    ct = fminf(fmaxf(ct, -1.0f), 1.0f);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float sp = __builtin_amdgcn_sinf(u2), cp = __builtin_amdgcn_cosf(u2);
    // orthonormal basis around mu (Duff et al. 2017); |sign + mu_z| >= 1
    float sign = copysignf(1.0f, l.mz);
    float a = -__builtin_amdgcn_rcpf(sign + l.mz);
    float b = l.mx * l.my * a;
    float tx = 1.0f + sign * l.mx * l.mx * a, ty = sign * b, tz = -sign * l.mx;
    float bx = b, by = sign + l.my * l.my * a, bz = -l.my;
    *wx = tx * st * cp + bx * st * sp + l.mx * ct;
    *wy = ty * st * cp + by * st * sp + l.my * ct;
    *wz = tz * st * cp + bz * st * sp + l.mz * ct;
}

}  // namespace trg
#endif
