// Path-guiding network: parameters, Adam training on the matrix cores, batched evaluation, and the
// per-wave training loop used by tr_render_transient_guided. Layouts are described in tr_guide.h.
//
// Training objective (neural importance sampling, Mueller et al. 2019): for a vertex whose next
// direction w was drawn from the one-sample MIS density p(w) = a p_bsdf(w) + (1 - a) q(w), the
// record weight is F(w) / p(w) with F = incident radiance x BSDF x cosine (luminance), and the loss
// is -weight * log q(w), the Monte-Carlo estimate of the KL divergence from F to q up to constants.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "tr_api.h"
#include "tr_device.h"
#include "tr_guide.h"
#include "tr_guide_host.h"
#include "tr_internal.h"

using namespace trg;

#define GHIP_TRY(expr)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return tri::fail(TR_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

namespace {

// waits for a stream's queued work when it goes out of scope (asynchronous uploads from host
// buffers that live on the caller's stack)
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

constexpr int kRecFloats = 16;  // training record: 4 float4 (see GuideView / record layout below)
// record layout: [0] pos (normalised) xyz, t/total_time; [1] wo xyz, medium flag;
//                [2] normal xyz, weight (F/p before normalisation); [3] sampled direction xyz, 0

// ---- parameter packing ------------------------------------------------------------------------
// forward fragment (layer l, out tile nt, input pair kp), lane, element j:
//   W_l[kp*32 + kperm(g, j)][nt*16 + lane%16], g = lane/16
// backward fragment (layer l = 1..4, input tile it, output pair op), lane, element j:
//   W_l[it*16 + lane%16][op*32 + kperm(g, j)]
__global__ void pack_kernel(const float* __restrict__ master, uint32_t* __restrict__ fwd, uint32_t* __restrict__ bwd,
                            float* __restrict__ bias)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_fwd = kFwdFrags * 64 * 4, n_bwd = kBwdFrags * 64 * 4;  // uint32 words (2 bf16 each)
    if (tid < n_fwd) {
        int word = tid & 3, lane = (tid >> 2) & 63, f = tid >> 8;
        int l = 0;
        while (l + 1 < kLayers && f >= fwd_frag_offset(l + 1)) ++l;
        int local = f - fwd_frag_offset(l), nt = local >> 1, kp = local & 1;
        int g = lane >> 4, o = nt * 16 + (lane & 15);
        const float* W = master + param_offset(l);
        float v[2];
        for (int h = 0; h < 2; ++h) {
            int i = kp * 32 + kperm(g, 2 * word + h);
            v[h] = (i < kIn[l] && o < kOutL[l]) ? W[i * kOutL[l] + o] : 0.0f;
        }
        fwd[tid] = pack_bf16(v[0], v[1]);
    } else if (tid < n_fwd + n_bwd) {
        int t = tid - n_fwd;
        int word = t & 3, lane = (t >> 2) & 63, f = t >> 8;
        int l = 1 + f / kBwdFragsPerLayer, local = f % kBwdFragsPerLayer, it = local >> 1, op = local & 1;
        int g = lane >> 4, i = it * 16 + (lane & 15);
        const float* W = master + param_offset(l);
        float v[2];
        for (int h = 0; h < 2; ++h) {
            int o = op * 32 + kperm(g, 2 * word + h);
            v[h] = (i < kIn[l] && o < kOutL[l]) ? W[i * kOutL[l] + o] : 0.0f;
        }
        bwd[t] = pack_bf16(v[0], v[1]);
    } else if (tid < n_fwd + n_bwd + kBiasPad) {
        int t = tid - n_fwd - n_bwd, l = t >> 6, o = t & 63;
        bias[t] = o < kOutL[l] ? master[param_offset(l) + kIn[l] * kOutL[l] + o] : 0.0f;
    }
}

// ---- batched evaluation -----------------------------------------------------------------------
constexpr int kEvalBlock = 256;  // 4 wavefronts, 64 samples each per iteration

template <int NST>
__global__ void __launch_bounds__(kEvalBlock) eval_kernel(const float* __restrict__ raw, uint32_t n,
                                                          const uint4* __restrict__ frag,
                                                          const float* __restrict__ bias, float* __restrict__ out)
{
    // features (bf16 rows) then outputs per wavefront, and the weight fragments: 80.9 KB, two
    // blocks per CU (the biases are read from global memory, where they stay in the L1 cache)
    __shared__ float lds[kEvalBlock / 64][64 * kRow];
    __shared__ uint4 wfrag[kFwdFrags * 64];
    for (int k = threadIdx.x; k < kFwdFrags * 64; k += kEvalBlock) wfrag[k] = frag[k];
    __syncthreads();
    const int lane = int(__lane_id()), wid = int(threadIdx.x >> 6);
    float* mine = lds[wid];
    const uint32_t waves = gridDim.x * (kEvalBlock / 64);
    for (uint32_t base = (blockIdx.x * (kEvalBlock / 64) + wid) * 64u; base < n; base += waves * 64u) {
        const uint32_t s = base + uint32_t(lane);
        float r[kRaw] = {};
        if (s < n)
            for (int k = 0; k < kRaw; ++k) r[k] = raw[size_t(s) * kRaw + k];
        float f[kFeat];
        encode_features(r, f);
        store_feature_row(reinterpret_cast<uint32_t*>(mine) + lane * kRow, f);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        mlp_forward_wave<NST>(mine, wfrag, bias, nullptr);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (s < n)
            for (int o = 0; o < kOut; ++o) out[size_t(s) * kOut + o] = mine[lane * kRow + o];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---- loss gradient of one record with respect to the 40 raw outputs ---------------------------
// Returns the loss -w log q(omega); d[] receives dL/d(raw).
__device__ float loss_grad(const float* o, float wx, float wy, float wz, float w, float* d)
{
    float logit_max = -3.4e38f;
    for (int i = 0; i < kLobes; ++i) logit_max = fmaxf(logit_max, o[5 * i + 4]);
    float se = 0.f;
    for (int i = 0; i < kLobes; ++i) se += __expf(o[5 * i + 4] - logit_max);
    const float log_se = logf(se);
    float lt[kLobes], mu_dot[kLobes], kap[kLobes], rn[kLobes], mu[kLobes][3];
    float lmax = -3.4e38f;
    for (int i = 0; i < kLobes; ++i) {
        float x = o[5 * i], y = o[5 * i + 1], z = o[5 * i + 2];
        float r = sqrtf(x * x + y * y + z * z);
        rn[i] = r;
        if (r > 1e-20f) {
            mu[i][0] = x / r, mu[i][1] = y / r, mu[i][2] = z / r;
        } else {
            mu[i][0] = 0.f, mu[i][1] = 0.f, mu[i][2] = 1.f;
        }
        float k = fminf(softplus(o[5 * i + 3]) + kKappaMin, kKappaMax);
        kap[i] = k;
        float c = mu[i][0] * wx + mu[i][1] * wy + mu[i][2] * wz;
        mu_dot[i] = c;
        // log pi_i + log vMF_i(omega)
        float log_pi = o[5 * i + 4] - logit_max - log_se;
        float log_norm = logf(k) - logf(2.0f * 3.14159265358979f) - logf(-expm1f(-2.0f * k));
        lt[i] = log_pi + log_norm + k * (c - 1.0f);
        lmax = fmaxf(lmax, lt[i]);
    }
    float sl = 0.f;
    for (int i = 0; i < kLobes; ++i) sl += __expf(lt[i] - lmax);
    const float log_q = lmax + logf(sl);
    for (int i = 0; i < kLobes; ++i) {
        float gamma = __expf(lt[i] - log_q);        // posterior responsibility of lobe i
        float pi = __expf(o[5 * i + 4] - logit_max - log_se);
        float dlogf = -w * gamma;                   // dL / d log vMF_i
        d[5 * i + 4] = -w * (gamma - pi);           // softmax logits
        float k = kap[i];
        // d log vMF / d kappa = 1/k - 2/(e^{2k} - 1) + (mu.w - 1); series for small kappa
        float gk = k < 1e-3f ? 1.0f - k / 3.0f : 1.0f / k - 2.0f / expm1f(2.0f * k);
        float dk = dlogf * (gk + mu_dot[i] - 1.0f);
        float raw_k = o[5 * i + 3];
        float sig = 1.0f / (1.0f + __expf(-raw_k));
        bool clamped = softplus(raw_k) + kKappaMin >= kKappaMax;
        d[5 * i + 3] = clamped ? 0.0f : dk * sig;
        // d log vMF / d mu = kappa * omega; mu = r / |r|
        if (rn[i] > 1e-20f) {
            float s = dlogf * k / rn[i];
            float c = mu_dot[i];
            d[5 * i + 0] = s * (wx - c * mu[i][0]);
            d[5 * i + 1] = s * (wy - c * mu[i][1]);
            d[5 * i + 2] = s * (wz - c * mu[i][2]);
        } else {
            d[5 * i + 0] = d[5 * i + 1] = d[5 * i + 2] = 0.f;
        }
    }
    return -w * log_q;
}

// ---- training: tiles of 64 records, four wavefronts per tile -----------------------------------
// LDS: acts[5][64][64] bf16 neuron-major layer inputs (40 KiB), U: feature rows / raw outputs
// (10.25 KiB), and dY[64][64] bf16 neuron-major (8 KiB) in U's memory once the outputs are read:
// 50.3 KiB per block, three blocks per CU.
// A tile's work is split over the block's four wavefronts: wavefront w evaluates sample tile w of
// the forward pass, owns output tile w of each layer's weight gradient, outputs 16 w .. 16 w + 15
// of its bias gradient and input tile w of its input gradient; the per-record work (features, loss
// gradient) runs on 16 lanes of each wavefront, records 16 w .. 16 w + 15. With one wavefront per
// tile the kernel was bound by that wavefront's instruction issue (22,500 instructions per tile,
// 10,000 of them scalar, one issue slot per 4 cycles: 85 us per step over 16,384 records, one
// wavefront per CU).
// Block b takes tiles b, b + gridDim.x, ... of the step's records and writes their gradient (sum
// over the records) to partial[b][kParams] and their loss sum to partial_loss[b]; the block's first
// tile stores, later ones add (each lane reads back only what it wrote: the split above is the same
// for every tile). Launches use one block per tile (train_grid). The weight-gradient tiles come out
// of the matrix cores row-major in the parameter layout (W_l[i][o], lanes of a 16-lane group on
// consecutive o), so every store writes 64-byte segments.
constexpr int kTrainWaves = 4;
constexpr int kTrainBlock = 64 * kTrainWaves;

struct TrainArgs {
    const float4* rec;
    const uint32_t* count;   // records available (device)
    const float* wsum;       // sum of record weights (device), for normalisation
    uint32_t first;          // first record of this step
    uint32_t batch;          // records per step
    const uint4* frag;
    const uint4* bfrag;
    const float* bias;
    float* partial;
    float* partial_loss;
};

// One 64-record tile of a training step: forward, loss gradient, backward; adds the tile's weight
// and bias gradient into the block's partial P (stores when `add` is false) and returns this
// wavefront's share of the tile's loss.
__device__ __forceinline__ float train_tile(const TrainArgs& A, uint32_t r0, uint32_t total, uint32_t avail, bool add,
                                            float* P, uint16_t (*acts)[64 * 64], uint16_t* dY, float* U)
{
    const int lane = int(__lane_id()), g = lane >> 4, c = lane & 15, wid = int(threadIdx.x >> 6);
    // this lane's record (lanes 0..15 of each wavefront): record j = first + ridx of the step
    // sequence is read at (j * p) mod total, p = 2^31 - 1 (a prime, so the map is a bijection of the
    // total records): each step's batch is spread over the whole pass instead of a run of adjacent
    // paths (neighbouring pixels), whatever order the batch was built in
    const bool rec_lane = lane < 16;
    const int rl = wid * 16 + c;
    const uint32_t ridx = r0 + uint32_t(rl);
    const bool valid = rec_lane && ridx < avail;
    float4 q0 = {}, q1 = {}, q2 = {}, q3 = {};
    if (valid) {
        const uint32_t src = uint32_t((uint64_t(A.first + ridx) * 2147483647ull) % uint64_t(total));
        const float4* r = A.rec + size_t(src) * 4;
        q0 = r[0], q1 = r[1], q2 = r[2], q3 = r[3];
    }
    const float mean_w = *A.wsum / float(max(total, 1u));
    const float w = valid && mean_w > 0.f ? q2.w / mean_w : 0.0f;
    // features: sample-major rows for the forward pass, neuron-major copy for the weight gradient
    if (rec_lane) {
        float raw[kRaw] = {q0.x, q0.y, q0.z, q1.x, q1.y, q1.z, q2.x, q2.y, q2.z, q0.w, q1.w, 0.f};
        float f[kFeat];
        encode_features(raw, f);
        store_feature_row(reinterpret_cast<uint32_t*>(U) + rl * kRow, f);
#pragma unroll
        for (int k = 0; k < 64; ++k) acts[0][k * 64 + rl] = k < kFeat ? uint16_t(bf16_bits(f[k])) : uint16_t(0);
    }
    __syncthreads();
    mlp_forward_wave<1>(U, A.frag, A.bias, &acts[1][0], wid, wid + 1);
    __syncthreads();
    // output gradient, one record per lane
    float loss = 0.f;
    {
        float o[kOut], d[kOut];
        if (rec_lane) {
#pragma unroll
            for (int k = 0; k < kOut; ++k) o[k] = U[rl * kRow + k];
        }
        __syncthreads();  // every output read before dY (the same LDS) is written
        if (rec_lane) {
            if (w > 0.f) {
                loss = loss_grad(o, q3.x, q3.y, q3.z, w, d);
            } else {
#pragma unroll
                for (int k = 0; k < kOut; ++k) d[k] = 0.f;
            }
#pragma unroll
            for (int k = 0; k < 64; ++k) dY[k * 64 + rl] = k < kOut ? uint16_t(bf16_bits(d[k])) : uint16_t(0);
        }
        // wavefront loss sum
        for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    }
    __syncthreads();
    for (int l = kLayers - 1; l >= 0; --l) {
        const int n_in = kIn[l], n_out = kOutL[l];
        const uint16_t* X = acts[l];
        // weight gradient D[i][o] = sum_s X[i][s] dY[o][s] of output tile ot = wid, stored as
        // P[W_l + i*n_out + o]: the output tile's rows are inputs i (lane group g, element j), its
        // columns outputs o (lane c)
        if (wid * 16 < n_out) {
            const int ot = wid;
            for (int it = 0; it < n_in / 16; ++it) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    uint4 a = *reinterpret_cast<const uint4*>(X + (it * 16 + c) * 64 + ks * 32 + 8 * g);
                    uint4 b = *reinterpret_cast<const uint4*>(dY + (ot * 16 + c) * 64 + ks * 32 + 8 * g);
                    acc = mfma32(a, b, acc);
                }
                const int o = ot * 16 + c;
                if (o < n_out)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float* q = P + param_offset(l) + (it * 16 + 4 * g + j) * n_out + o;
                        *q = add ? *q + acc[j] : acc[j];
                    }
            }
        }
        // bias gradient of outputs 16 wid + c: lane group g sums samples 16 g .. 16 g + 15
        {
            const int o = wid * 16 + c;
            float sb = 0.f;
            if (o < n_out)
#pragma unroll
                for (int s = 0; s < 16; ++s) sb += bf16_to_f32(dY[o * 64 + g * 16 + s]);
            sb += __shfl_xor(sb, 16);
            sb += __shfl_xor(sb, 32);
            if (g == 0 && o < n_out) {
                float* q = P + param_offset(l) + n_in * n_out + o;
                *q = add ? *q + sb : sb;
            }
        }
        if (l == 0) break;
        // input gradient dX[i][s] = sum_o W[i][o] dY[o][s] of input tile it = wid, masked by
        // ReLU'(layer input)
        f32x4 acc[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) acc[st] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int op = 0; op < 2; ++op) {
            uint4 B[4];
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                uint32_t wv[4];
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    uint32_t lo = dY[(op * 32 + kperm(g, 2 * h)) * 64 + st * 16 + c];
                    uint32_t hi = dY[(op * 32 + kperm(g, 2 * h + 1)) * 64 + st * 16 + c];
                    wv[h] = lo | (hi << 16);
                }
                B[st] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
            }
            uint4 a = A.bfrag[((l - 1) * kBwdFragsPerLayer + wid * 2 + op) * 64 + lane];
#pragma unroll
            for (int st = 0; st < 4; ++st) acc[st] = mfma32(a, B[st], acc[st]);
        }
        __syncthreads();  // every wavefront has read this layer's dY
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int i = wid * 16 + 4 * g + j, s = st * 16 + c;
                float v = X[i * 64 + s] != 0 ? acc[st][j] : 0.0f;  // ReLU' of the layer input
                dY[i * 64 + s] = uint16_t(bf16_bits(v));
            }
        __syncthreads();
    }
    return loss;
}

__global__ void __launch_bounds__(kTrainBlock) train_kernel(TrainArgs A)
{
    __shared__ uint16_t acts[kLayers][64 * 64];
    // U (feature rows, then the raw outputs) and dY share their LDS: the outputs are read into
    // registers before the first output gradient is written
    __shared__ float U[64 * kRow];
    __shared__ float wave_loss[kTrainWaves];
    static_assert(64 * 64 * sizeof(uint16_t) <= sizeof(U), "dY fits in U");
    uint16_t* dY = reinterpret_cast<uint16_t*>(U);
    const uint32_t total = *A.count;
    const uint32_t avail = total > A.first ? min(total - A.first, A.batch) : 0u;
    float* P = A.partial + size_t(blockIdx.x) * kParams;
    float loss = 0.f;
    for (uint32_t r0 = blockIdx.x * 64u; r0 < avail; r0 += gridDim.x * 64u) {
        loss += train_tile(A, r0, total, avail, r0 != blockIdx.x * 64u, P, acts, dY, U);
        __syncthreads();  // the next tile overwrites the LDS arrays
    }
    if (__lane_id() == 0) wave_loss[threadIdx.x >> 6] = loss;
    __syncthreads();
    if (threadIdx.x == 0 && blockIdx.x * 64u < avail) {
        float s = 0.f;
        for (int k = 0; k < kTrainWaves; ++k) s += wave_loss[k];
        A.partial_loss[blockIdx.x] = s;
    }
}

// Blocks of a training step: one per tile of 64 records
inline uint32_t train_grid(uint32_t records)
{
    return std::max(1u, (records + 63u) / 64u);
}

// First stage of the gradient sum: block row y adds the partials of its slice of the training
// grid's blocks (those that had a tile), in block order, so the sum is deterministic and the Adam
// kernel adds only kReduceSlices values per parameter.
constexpr uint32_t kReduceSlices = 32;

__global__ void reduce_kernel(const float* __restrict__ partial, const float* __restrict__ partial_loss,
                              const uint32_t* count, uint32_t first, uint32_t batch, uint32_t train_blocks,
                              float* __restrict__ part2, float* __restrict__ loss2)
{
    const uint32_t total = *count;
    const uint32_t avail = total > first ? min(total - first, batch) : 0u;
    const uint32_t chunks = min((avail + 63u) / 64u, train_blocks);
    const uint32_t per = (chunks + kReduceSlices - 1u) / kReduceSlices;
    const uint32_t c0 = min(blockIdx.y * per, chunks), c1 = min(c0 + per, chunks);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float ls = 0.f;
        for (uint32_t c = c0; c < c1; ++c) ls += partial_loss[c];
        loss2[blockIdx.y] = ls;
    }
    if (p >= kParams) return;
    float sum = 0.f;
    for (uint32_t c = c0; c < c1; ++c) sum += partial[size_t(c) * kParams + p];
    part2[size_t(blockIdx.y) * kParams + p] = sum;
}

// Adam (Kingma & Ba 2015) on the master FP32 parameters: gradient = mean over the step's records.
struct AdamArgs {
    float* master;
    float* m;
    float* v;
    const float* partial;       // [kReduceSlices][kParams] slice sums (reduce_kernel)
    const float* partial_loss;  // [kReduceSlices]
    const uint32_t* count;
    uint32_t first, batch;
    float lr, beta1, beta2, eps;
    float bc1, bc2;  // 1 - beta^t
    float* loss_out; // mean loss of the step (device)
};

__global__ void adam_kernel(AdamArgs A)
{
    const uint32_t total = *A.count;
    const uint32_t avail = total > A.first ? min(total - A.first, A.batch) : 0u;
    if (avail == 0) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        float ls = 0.f;
        for (uint32_t c = 0; c < kReduceSlices; ++c) ls += A.partial_loss[c];
        *A.loss_out = ls / float(avail);
    }
    if (p >= kParams) return;
    float gsum = 0.f;
    for (uint32_t c = 0; c < kReduceSlices; ++c) gsum += A.partial[size_t(c) * kParams + p];
    float gr = gsum / float(avail);
    float m = A.beta1 * A.m[p] + (1.0f - A.beta1) * gr;
    float v = A.beta2 * A.v[p] + (1.0f - A.beta2) * gr * gr;
    A.m[p] = m;
    A.v[p] = v;
    float mh = m / A.bc1, vh = v / A.bc2;
    A.master[p] -= A.lr * mh / (sqrtf(vh) + A.eps);
}

// After a wave: records of training paths -> batch (weight = pre * sum of later contributions).
// One thread per training path walks its levels from the last down, keeping the running sum of the
// later levels' contributions, in two passes: the first counts the records it keeps (one block
// atomic on the batch count per 256 paths), the second writes them; the weight sum is added once
// per block. (A thread per record that re-summed the later levels with dependent loads took
// 445 us per wave over 2^19 paths x 17 levels.) The batch's order is not the order the training
// steps see: they visit it through a bijection (train_tile).
__global__ void build_batch_kernel(trg::GuideView V, float4* __restrict__ batch, uint32_t* __restrict__ count,
                                   float* __restrict__ wsum)
{
    __shared__ uint32_t scan[256 / 64 + 1];
    __shared__ float wred[256 / 64];
    const uint32_t L = V.max_levels;
    float ws = 0.f;
    for (uint32_t base = blockIdx.x * blockDim.x; base < V.train_paths; base += gridDim.x * blockDim.x) {
        const uint32_t tp = base + threadIdx.x;
        const uint32_t nr = tp < V.train_paths ? min(V.n_rec[tp], L) : 0u;
        const float* cb = V.contrib + size_t(tp) * L;
        const float4* rb = V.rec + size_t(tp) * L * 4;
        uint32_t kept = 0;
        if (nr > 0) {
            float run = 0.f;
            for (int lvl = int(L) - 1; lvl >= 0; --lvl) {
                if (uint32_t(lvl) < nr) {
                    const float w = rb[size_t(lvl) * 4 + 2].w * run;
                    kept += (w > 0.f && isfinite(w)) ? 1u : 0u;
                }
                run += cb[lvl];
            }
        }
        uint32_t pos = trd::block_append(count, kept, scan);
        if (kept > 0) {
            float run = 0.f;
            for (int lvl = int(L) - 1; lvl >= 0; --lvl) {
                if (uint32_t(lvl) < nr) {
                    const float4* r = rb + size_t(lvl) * 4;
                    float4 r2 = r[2];
                    const float w = r2.w * run;
                    if (w > 0.f && isfinite(w)) {
                        r2.w = w;
                        float4* b = batch + size_t(pos) * 4;
                        b[0] = r[0], b[1] = r[1], b[2] = r2, b[3] = r[3];
                        ++pos;
                        ws += w;
                    }
                }
                run += cb[lvl];
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) ws += __shfl_xor(ws, off);
    if (__lane_id() == 0) wred[threadIdx.x >> 6] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (uint32_t k = 0; k < blockDim.x / 64; ++k) t += wred[k];
        if (t > 0.f) atomicAdd(wsum, t);
    }
}

__global__ void clear_kernel(uint32_t* __restrict__ a, size_t na, float* __restrict__ b, size_t nb,
                             uint32_t* __restrict__ c, size_t nc)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < na + nb + nc; i += stride) {
        if (i < na) a[i] = 0u;
        else if (i < na + nb) b[i - na] = 0.f;
        else c[i - na - nb] = 0u;
    }
}

template <class T>
int galloc(T** p, size_t n)
{
    *p = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return tri::fail(TR_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return TR_OK;
}
template <class T>
void gfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// deterministic initial weights: Glorot-uniform from a PCG32 stream; the output layer is scaled
// down so that the untrained mixture is close to uniform on the sphere
std::vector<float> initial_params(uint64_t seed)
{
    std::vector<float> p(kParams, 0.0f);
    uint64_t state = 0, inc = (0x9e3779b97f4a7c15ULL << 1) | 1u;
    auto next = [&]() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u), rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    };
    next();
    state += seed;
    next();
    for (int l = 0; l < kLayers; ++l) {
        float b = std::sqrt(6.0f / float(kIn[l] + kOutL[l]));
        if (l == kLayers - 1) b *= 0.1f;
        float* W = p.data() + param_offset(l);
        for (int k = 0; k < kIn[l] * kOutL[l]; ++k) W[k] = (float(next() >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * b;
    }
    return p;
}

}  // namespace

struct tr_guide_s {
    tr_guide_desc desc{};
    hipStream_t stream = nullptr;
    float *master = nullptr, *m = nullptr, *v = nullptr, *bias = nullptr;
    uint32_t *fwd = nullptr, *bwd = nullptr;
    float *partial = nullptr, *partial_loss = nullptr;
    float *part2 = nullptr, *loss2 = nullptr;  // slice sums of the per-chunk gradients
    uint32_t max_chunks = 0;
    // render-time buffers
    float4* rec = nullptr;
    uint32_t* n_rec = nullptr;
    float* contrib = nullptr;
    float4* batch = nullptr;
    uint32_t* count = nullptr;  // [0] batch count
    float* wsum = nullptr;      // [0] weight sum, [1] last loss
    size_t rec_cap = 0, tp_cap = 0;
    trg::GuideView view{};
    uint64_t steps = 0, records = 0;
    double train_ms = 0.0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int eval_tiles = 2;  // sample tiles per pass of the batched evaluation kernel

    void release_render()
    {
        gfree(rec), gfree(n_rec), gfree(contrib), gfree(batch);
        rec_cap = tp_cap = 0;
    }
    ~tr_guide_s()
    {
        release_render();
        gfree(master), gfree(m), gfree(v), gfree(bias), gfree(fwd), gfree(bwd), gfree(partial), gfree(partial_loss);
        gfree(part2), gfree(loss2), gfree(count), gfree(wsum);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int pack(hipStream_t s)
    {
        const int n = kFwdFrags * 256 + kBwdFrags * 256 + kBiasPad;
        pack_kernel<<<(n + 255) / 256, 256, 0, s>>>(master, fwd, bwd, bias);
        GHIP_TRY(hipGetLastError());
        return TR_OK;
    }
    // Adam steps over the batch in device memory (count and weight sum on the device)
    int train(hipStream_t s, uint32_t steps_now)
    {
        for (uint32_t k = 0; k < steps_now; ++k) {
            TrainArgs T;
            T.rec = batch ? batch : nullptr;
            T.count = count;
            T.wsum = wsum;
            T.first = k * desc.batch_size;
            T.batch = desc.batch_size;
            T.frag = reinterpret_cast<const uint4*>(fwd);
            T.bfrag = reinterpret_cast<const uint4*>(bwd);
            T.bias = bias;
            T.partial = partial;
            T.partial_loss = partial_loss;
            const uint32_t tg = train_grid(desc.batch_size);
            train_kernel<<<tg, kTrainBlock, 0, s>>>(T);
            GHIP_TRY(hipGetLastError());
            ++steps;
            AdamArgs A;
            reduce_kernel<<<dim3((kParams + 255) / 256, kReduceSlices), 256, 0, s>>>(partial, partial_loss, count, T.first,
                                                                                     T.batch, tg, part2, loss2);
            GHIP_TRY(hipGetLastError());
            A.master = master, A.m = m, A.v = v, A.partial = part2, A.partial_loss = loss2;
            A.count = count, A.first = T.first, A.batch = T.batch;
            A.lr = desc.learning_rate, A.beta1 = 0.9f, A.beta2 = 0.999f, A.eps = 1e-8f;
            A.bc1 = float(1.0 - std::pow(0.9, double(steps)));
            A.bc2 = float(1.0 - std::pow(0.999, double(steps)));
            A.loss_out = wsum + 1;
            adam_kernel<<<(kParams + 255) / 256, 256, 0, s>>>(A);
            GHIP_TRY(hipGetLastError());
            int rc = pack(s);
            if (rc) return rc;
        }
        return TR_OK;
    }
};

namespace trg {

int guide_begin_render(tr_guide g, uint32_t wave, uint32_t max_levels, const float bb_lo[3], const float bb_hi[3],
                       hipStream_t stream, GuideView* view)
{
    (void)stream;
    // a frozen network (no Adam steps) collects no training records
    uint32_t tp = g->desc.steps_per_wave ? std::min<uint32_t>(g->desc.train_paths, wave) : 0u;
    size_t nrec = size_t(tp) * max_levels;
    if (tp > g->tp_cap || nrec > g->rec_cap) {
        g->release_render();
        int rc;
        if ((rc = galloc(&g->rec, nrec * 4)) || (rc = galloc(&g->n_rec, tp)) || (rc = galloc(&g->contrib, nrec)) ||
            (rc = galloc(&g->batch, nrec * 4)))
            return rc;
        g->tp_cap = tp;
        g->rec_cap = nrec;
    }
    GuideView& V = g->view;
    V.frag = reinterpret_cast<const uint4*>(g->fwd);
    V.bias = g->bias;
    V.rec = g->rec;
    V.n_rec = g->n_rec;
    V.contrib = g->contrib;
    V.train_paths = tp;
    V.stride = tp ? std::max<uint32_t>(1u, wave / tp) : 0xFFFFFFFFu;
    V.max_levels = max_levels;
    V.alpha = g->desc.bsdf_fraction;
    float ext[3];
    for (int k = 0; k < 3; ++k) ext[k] = std::max(bb_hi[k] - bb_lo[k], 1e-6f);
    V.bb_lo = make_float3(bb_lo[0], bb_lo[1], bb_lo[2]);
    V.bb_inv = make_float3(1.0f / ext[0], 1.0f / ext[1], 1.0f / ext[2]);
    *view = V;
    return TR_OK;
}

int guide_wave_begin(tr_guide g, hipStream_t s)
{
    if (g->view.train_paths == 0) return TR_OK;
    size_t na = g->view.train_paths, nb = size_t(g->view.train_paths) * g->view.max_levels;
    clear_kernel<<<256, 256, 0, s>>>(g->n_rec, na, g->contrib, nb, g->count, 1);
    GHIP_TRY(hipGetLastError());
    clear_kernel<<<1, 64, 0, s>>>(nullptr, 0, g->wsum, 2, nullptr, 0);
    GHIP_TRY(hipGetLastError());
    return TR_OK;
}

int guide_wave_end(tr_guide g, hipStream_t s)
{
    if (g->view.train_paths == 0 || g->desc.steps_per_wave == 0) return TR_OK;
    uint32_t grid = std::max(1u, std::min((g->view.train_paths + 255u) / 256u, 4096u));
    build_batch_kernel<<<grid, 256, 0, s>>>(g->view, g->batch, g->count, g->wsum);
    GHIP_TRY(hipGetLastError());
    return g->train(s, g->desc.steps_per_wave);
}

}  // namespace trg

extern "C" {

int tr_guide_create(tr_guide* out, const tr_guide_desc* d)
{
    if (!out) return tri::fail(TR_ERR_INVALID, "tr_guide_create: null output");
    *out = nullptr;
    tr_guide_desc D{};
    D.learning_rate = 1e-3f;
    D.bsdf_fraction = 0.5f;
    D.train_paths = 1u << 16;
    D.batch_size = 1u << 14;
    D.steps_per_wave = 1;
    D.init_seed = 1;
    if (d) D = *d;
    if (D.batch_size == 0 || D.batch_size % 64 != 0 || D.batch_size > (1u << 20))
        return tri::fail(TR_ERR_INVALID, "tr_guide_create: batch_size must be a multiple of 64 in (0, 2^20]");
    if (!(D.bsdf_fraction > 0.0f && D.bsdf_fraction <= 1.0f))
        return tri::fail(TR_ERR_INVALID, "tr_guide_create: bsdf_fraction must be in (0, 1] (MIS needs the BSDF)");
    if (!(D.learning_rate >= 0.0f)) return tri::fail(TR_ERR_INVALID, "tr_guide_create: negative learning rate");
    int rc = tri::check_device();
    if (rc) return rc;
    tr_guide g = new (std::nothrow) tr_guide_s;
    if (!g) return tri::fail(TR_ERR_NOMEM, "tr_guide_create: out of host memory");
    g->desc = D;
    g->max_chunks = train_grid(D.batch_size);  // partials: one per training block
    auto bail = [&](int code) {
        delete g;
        return code;
    };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(tri::fail(TR_ERR_DEVICE, "tr_guide_create: stream"));
    for (auto& e : g->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(tri::fail(TR_ERR_DEVICE, "tr_guide_create: event"));
    if ((rc = galloc(&g->master, kParams)) || (rc = galloc(&g->m, kParams)) || (rc = galloc(&g->v, kParams)) ||
        (rc = galloc(&g->bias, kBiasPad)) || (rc = galloc(&g->fwd, size_t(kFwdFrags) * 256)) ||
        (rc = galloc(&g->bwd, size_t(kBwdFrags) * 256)) ||
        (rc = galloc(&g->partial, size_t(g->max_chunks) * kParams)) || (rc = galloc(&g->partial_loss, g->max_chunks)) ||
        (rc = galloc(&g->part2, size_t(kReduceSlices) * kParams)) || (rc = galloc(&g->loss2, kReduceSlices)) ||
        (rc = galloc(&g->count, 4)) || (rc = galloc(&g->wsum, 4)))
        return bail(rc);
    std::vector<float> p = initial_params(D.init_seed);
    if ((rc = tr_guide_set_params(g, p.data()))) return bail(rc);
    *out = g;
    return TR_OK;
}

int tr_guide_set_steps(tr_guide g, uint32_t steps_per_wave)
{
    if (!g) return tri::fail(TR_ERR_INVALID, "tr_guide_set_steps: null guide");
    g->desc.steps_per_wave = steps_per_wave;
    return TR_OK;
}

int tr_guide_destroy(tr_guide g)
{
    if (g) {
        (void)hipStreamSynchronize(g->stream);
        delete g;
    }
    return TR_OK;
}

int tr_guide_get_params(tr_guide g, float* params)
{
    if (!g || !params) return tri::fail(TR_ERR_INVALID, "tr_guide_get_params: null argument");
    GHIP_TRY(hipMemcpyAsync(params, g->master, kParams * sizeof(float), hipMemcpyDeviceToHost, g->stream));
    GHIP_TRY(hipStreamSynchronize(g->stream));
    return TR_OK;
}

int tr_guide_set_params(tr_guide g, const float* params)
{
    if (!g || !params) return tri::fail(TR_ERR_INVALID, "tr_guide_set_params: null argument");
    for (int i = 0; i < kParams; ++i)
        if (!std::isfinite(params[i])) return tri::fail(TR_ERR_INVALID, "tr_guide_set_params: non-finite parameter");
    GHIP_TRY(hipMemcpyAsync(g->master, params, kParams * sizeof(float), hipMemcpyHostToDevice, g->stream));
    clear_kernel<<<64, 256, 0, g->stream>>>(nullptr, 0, g->m, kParams, nullptr, 0);
    clear_kernel<<<64, 256, 0, g->stream>>>(nullptr, 0, g->v, kParams, nullptr, 0);
    GHIP_TRY(hipGetLastError());
    int rc = g->pack(g->stream);
    if (rc) return rc;
    GHIP_TRY(hipStreamSynchronize(g->stream));
    g->steps = 0;
    return TR_OK;
}

int tr_guide_eval_device(tr_guide g, const float* d_raw, uint32_t n, float* d_out, float* kernel_ms)
{
    if (!g || (n && (!d_raw || !d_out))) return tri::fail(TR_ERR_INVALID, "tr_guide_eval_device: null argument");
    if (n == 0) return TR_OK;
    int cu = tri::compute_units();
    uint32_t waves_needed = (n + 63u) / 64u;
    uint32_t grid = std::max(1u, std::min<uint32_t>((waves_needed + 3u) / 4u, uint32_t(std::max(cu, 1)) * 8u));
    GHIP_TRY(hipEventRecord(g->ev[0], g->stream));
    if (g->eval_tiles == 4)
        eval_kernel<4><<<grid, kEvalBlock, 0, g->stream>>>(d_raw, n, reinterpret_cast<const uint4*>(g->fwd), g->bias, d_out);
    else if (g->eval_tiles == 2)
        eval_kernel<2><<<grid, kEvalBlock, 0, g->stream>>>(d_raw, n, reinterpret_cast<const uint4*>(g->fwd), g->bias, d_out);
    else
        eval_kernel<1><<<grid, kEvalBlock, 0, g->stream>>>(d_raw, n, reinterpret_cast<const uint4*>(g->fwd), g->bias, d_out);
    GHIP_TRY(hipGetLastError());
    GHIP_TRY(hipEventRecord(g->ev[1], g->stream));
    GHIP_TRY(hipEventSynchronize(g->ev[1]));
    if (kernel_ms) {
        float ms = 0.f;
        GHIP_TRY(hipEventElapsedTime(&ms, g->ev[0], g->ev[1]));
        *kernel_ms = ms;
    }
    return TR_OK;
}

int tr_guide_eval(tr_guide g, const float* raw, uint32_t n, float* out)
{
    if (!g || (n && (!raw || !out))) return tri::fail(TR_ERR_INVALID, "tr_guide_eval: null argument");
    if (n == 0) return TR_OK;
    float *d_raw = nullptr, *d_out = nullptr;
    int rc;
    if ((rc = galloc(&d_raw, size_t(n) * kRaw))) return rc;
    if ((rc = galloc(&d_out, size_t(n) * kOut))) {
        gfree(d_raw);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_raw, raw, size_t(n) * kRaw * sizeof(float), hipMemcpyHostToDevice, g->stream);
    if (e == hipSuccess) {
        rc = tr_guide_eval_device(g, d_raw, n, d_out, nullptr);
        if (!rc) {
            e = hipMemcpyAsync(out, d_out, size_t(n) * kOut * sizeof(float), hipMemcpyDeviceToHost, g->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
        }
    }
    gfree(d_raw);
    gfree(d_out);
    if (rc) return rc;
    if (e != hipSuccess) return tri::fail(TR_ERR_DEVICE, std::string("tr_guide_eval: ") + hipGetErrorString(e));
    return TR_OK;
}

int tr_guide_gradient(tr_guide g, const float* records, uint32_t n, float* grad, float* loss)
{
    if (!g || (n && !records) || !grad) return tri::fail(TR_ERR_INVALID, "tr_guide_gradient: null argument");
    if (n == 0 || n > g->desc.batch_size)
        return tri::fail(TR_ERR_INVALID, "tr_guide_gradient: record count must be in [1, batch_size]");
    float4* d_rec = nullptr;
    int rc;
    if ((rc = galloc(&d_rec, size_t(n) * 4))) return rc;
    // records carry their own weights; normalise by the mean weight as training does
    double ws = 0.0;
    for (uint32_t i = 0; i < n; ++i) ws += records[size_t(i) * kRecFloats + 11];
    uint32_t cnt[4] = {n, 0, 0, 0};
    float wsv[4] = {float(ws), 0.f, 0.f, 0.f};
    // the uploads read cnt / wsv / records asynchronously: every return waits for them first
    StreamDrain drain{g->stream};
    GHIP_TRY(hipMemcpyAsync(d_rec, records, size_t(n) * kRecFloats * sizeof(float), hipMemcpyHostToDevice, g->stream));
    GHIP_TRY(hipMemcpyAsync(g->count, cnt, sizeof cnt, hipMemcpyHostToDevice, g->stream));
    GHIP_TRY(hipMemcpyAsync(g->wsum, wsv, sizeof wsv, hipMemcpyHostToDevice, g->stream));
    TrainArgs T;
    T.rec = d_rec, T.count = g->count, T.wsum = g->wsum, T.first = 0, T.batch = g->desc.batch_size;
    T.frag = reinterpret_cast<const uint4*>(g->fwd);
    T.bfrag = reinterpret_cast<const uint4*>(g->bwd);
    T.bias = g->bias, T.partial = g->partial, T.partial_loss = g->partial_loss;
    const uint32_t tg = train_grid(n);
    train_kernel<<<tg, kTrainBlock, 0, g->stream>>>(T);
    hipError_t e = hipGetLastError();
    std::vector<float> part(size_t(tg) * kParams), pl(tg);
    if (e == hipSuccess)
        e = hipMemcpyAsync(part.data(), g->partial, part.size() * sizeof(float), hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pl.data(), g->partial_loss, pl.size() * sizeof(float), hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    gfree(d_rec);
    if (e != hipSuccess) return tri::fail(TR_ERR_DEVICE, std::string("tr_guide_gradient: ") + hipGetErrorString(e));
    for (int p = 0; p < kParams; ++p) {
        double s = 0.0;
        for (size_t c = 0; c < pl.size(); ++c) s += part[c * kParams + size_t(p)];
        grad[p] = float(s / double(n));
    }
    if (loss) {
        double s = 0.0;
        for (float x : pl) s += x;
        *loss = float(s / double(n));
    }
    return TR_OK;
}

int tr_guide_train_batch(tr_guide g, const float* records, uint32_t n, float* loss)
{
    if (!g || (n && !records)) return tri::fail(TR_ERR_INVALID, "tr_guide_train_batch: null argument");
    if (n == 0 || n > g->desc.batch_size)
        return tri::fail(TR_ERR_INVALID, "tr_guide_train_batch: record count must be in [1, batch_size]");
    if (!g->batch || g->rec_cap < n) {
        g->release_render();
        int rc;
        if ((rc = galloc(&g->batch, size_t(g->desc.batch_size) * 4))) return rc;
        if ((rc = galloc(&g->rec, 4)) || (rc = galloc(&g->n_rec, 1)) || (rc = galloc(&g->contrib, 1))) return rc;
        g->rec_cap = g->desc.batch_size;
        g->tp_cap = 0;
    }
    double ws = 0.0;
    for (uint32_t i = 0; i < n; ++i) ws += records[size_t(i) * kRecFloats + 11];
    uint32_t cnt[4] = {n, 0, 0, 0};
    float wsv[4] = {float(ws), 0.f, 0.f, 0.f};
    StreamDrain drain{g->stream};
    GHIP_TRY(hipMemcpyAsync(g->batch, records, size_t(n) * kRecFloats * sizeof(float), hipMemcpyHostToDevice, g->stream));
    GHIP_TRY(hipMemcpyAsync(g->count, cnt, sizeof cnt, hipMemcpyHostToDevice, g->stream));
    GHIP_TRY(hipMemcpyAsync(g->wsum, wsv, sizeof wsv, hipMemcpyHostToDevice, g->stream));
    int rc = g->train(g->stream, 1);
    if (rc) return rc;
    float lv[2];
    GHIP_TRY(hipMemcpyAsync(lv, g->wsum, sizeof lv, hipMemcpyDeviceToHost, g->stream));
    GHIP_TRY(hipStreamSynchronize(g->stream));
    if (loss) *loss = lv[1];
    g->records += n;
    return TR_OK;
}

int tr_guide_info_get(tr_guide g, tr_guide_info* info)
{
    if (!g || !info) return tri::fail(TR_ERR_INVALID, "tr_guide_info_get: null argument");
    float lv[2] = {0.f, 0.f};
    GHIP_TRY(hipMemcpyAsync(lv, g->wsum, sizeof lv, hipMemcpyDeviceToHost, g->stream));
    GHIP_TRY(hipStreamSynchronize(g->stream));
    info->steps = g->steps;
    info->records = g->records;
    info->last_loss = lv[1];
    info->n_params = uint32_t(kParams);
    info->flop_per_eval = uint32_t(kFlopForward);
    return TR_OK;
}

}  // extern "C"
