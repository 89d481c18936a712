// Device-side building blocks for gfx950: PCG32 streams, watertight BVH traversal, light and
// BSDF sampling. Included only by HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "tr_internal.h"

namespace trd {

// ---- float3 helpers ------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 mk(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 normalize(f3 a)
{
    float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
__device__ __forceinline__ float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// ---- PCG32 (the reference's Random::RNG, src/Utils/RandomNumbers.h:119-174, 83-98), one stream
// per (seed, sequence). The increment is (seq << 1) | 1 and is recomputed from the sequence id. --
__device__ __forceinline__ uint32_t pcg_next(uint64_t& state, uint64_t inc)
{
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = uint32_t(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}
__device__ __forceinline__ float pcg_float(uint64_t& state, uint64_t inc)
{
    uint32_t bits = (pcg_next(state, inc) >> 9) | 0x3f800000u;
    return __uint_as_float(bits) - 1.0f;
}
__device__ __forceinline__ uint64_t pcg_inc(uint64_t seq) { return (seq << 1u) | 1u; }
__device__ __forceinline__ uint64_t pcg_seed(uint64_t initstate, uint64_t seq)
{
    uint64_t inc = pcg_inc(seq), s = 0u;
    pcg_next(s, inc);
    s += initstate;
    pcg_next(s, inc);
    return s;
}

// ---- wave-aggregated queue append (one atomic per wavefront) -------------------------------
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred)
{
    uint64_t mask = __ballot(pred);
    uint32_t lane = __lane_id();
    uint32_t base = 0;
    if (mask != 0) {
        int leader = __ffsll((long long)mask) - 1;
        if (int(lane) == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
        base = __shfl(base, leader);
    }
    return base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

// ---- block-aggregated append to two queues with one atomic ----------------------------------
// As block_append for two queues whose lengths are the low and high halves of one 64-bit counter
// (`pair`, 8-byte aligned): each thread passes its items for queue A (0 or 1) and queue B (0 or 1)
// and gets its positions in both; the block issues a single 64-bit atomic. Per-block counts fit
// 16 bits (blocks of at most 1024 threads), so the two scans run packed in one 32-bit value.
__device__ __forceinline__ void block_append2(uint32_t* pair, uint32_t cnt_a, uint32_t cnt_b, uint32_t* smem,
                                              uint32_t* pos_a, uint32_t* pos_b)
{
    const uint32_t lane = __lane_id(), wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const uint32_t mine = cnt_a | (cnt_b << 16);
    uint32_t x = mine;  // inclusive scan within the wavefront, both halves at once
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) smem[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            uint32_t t = smem[w];
            smem[w] = total;  // exclusive prefix of wavefront totals
            total += t;
        }
        unsigned long long old = 0ull;
        if (total)
            old = atomicAdd(reinterpret_cast<unsigned long long*>(pair),
                            (unsigned long long)(total & 0xFFFFu) | ((unsigned long long)(total >> 16) << 32));
        smem[nw] = uint32_t(old);
        smem[nw + 1] = uint32_t(old >> 32);
    }
    __syncthreads();
    const uint32_t pre = smem[wid] + x - mine;
    *pos_a = smem[nw] + (pre & 0xFFFFu);
    *pos_b = smem[nw + 1] + (pre >> 16);
    __syncthreads();  // smem may be reused by the caller's next iteration
}

// ---- block-aggregated queue append ----------------------------------------------------------
// Every thread of the block calls this with its number of items (0 or more); the block issues one
// atomic on `counter` and each thread gets the first queue position of its items. `smem` holds at
// least blockDim.x / 64 + 1 words. Same-address atomics from every wavefront are the limiting
// rate at tens of millions of queue entries, hence one per block.
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, uint32_t cnt, uint32_t* smem)
{
    const uint32_t lane = __lane_id(), wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    uint32_t x = cnt;  // inclusive scan within the wavefront
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) smem[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (uint32_t w = 0; w < nw; ++w) {
            uint32_t t = smem[w];
            smem[w] = total;  // exclusive prefix of wavefront totals
            total += t;
        }
        smem[nw] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    uint32_t pos = smem[nw] + smem[wid] + x - cnt;
    __syncthreads();  // smem may be reused by the caller's next iteration
    return pos;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ---- ray/box and watertight ray/triangle -----------------------------------------------------
struct RayPre {
    f3 o, d, inv;
    f3 c;       // -(o * inv): slab offset of the quantized 4-wide BVH's node test
    float k;    // slab margin of that test per unit |inv|
    f3 o_rot;   // origin components in (kx, ky, kz) order
    int rot;   // = kx: offset of the (kx, ky, kz) run in a rotated triangle vertex (TriRot)
    float Sx, Sy, Sz;
};

// reciprocal for slab tests; zero components become +-1e20 so that (lo - o) * inv never
// evaluates 0 * inf (a ray lying in a box face plane stays inside that slab)
__device__ __forceinline__ float safe_rcp(float d)
{
    return 1.0f / (fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d);
}

// Offset and margin of the quantized node test. A plane at P = origin + q * step (one rounding
// on the host) is entered at (P - o) * inv; the node test evaluates it as fma(q, step * inv, b)
// with b = fma(origin, inv, c), c = -(o * inv), widened to b -+ k |inv| (one fused multiply-add
// each; entry planes down, exit planes up), so a plane costs one operation and an axis three more
// per node. The roundings, each at most u = 2^-24 of its result, are: the host plane
// u (|origin| + 255 step) |inv|, o * inv u |o| |inv| (the negation is exact), b u |b| with
// |b| <= (|origin| + |o|) |inv| (1 + u), b -+ m u (|b| + m) and the plane distance
// u (|b| + 255 step |inv| + m), m = k |inv|. With C >= |origin| + 255 step for every node
// (plane_bound) they sum to at most u |inv| (4 C + 4 |o|) + 2 u m, which m exceeds for
// k = 6e-7 (C + max |o|) (about 10 u (C + |o|)): the entry plane is never late and the exit plane
// never early.
__device__ __forceinline__ RayPre ray_pre(f3 o, f3 d, float plane_bound)
{
    RayPre r;
    r.o = o;
    r.d = d;
    r.inv = mk(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    r.c = mk(-(o.x * r.inv.x), -(o.y * r.inv.y), -(o.z * r.inv.z));
    r.k = 6e-7f * (plane_bound + fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)));
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    // The paper swaps kx and ky when d[kz] < 0 to keep the winding; that only negates U, V, W, det
    // and T (each exactly), so the hit test and T / det are the same without it.
    r.rot = kx;
    r.o_rot = mk(comp(o, kx), comp(o, ky), comp(o, kz));
    r.Sz = 1.0f / comp(d, kz);
    r.Sx = comp(d, kx) * r.Sz;
    r.Sy = comp(d, ky) * r.Sz;
    return r;
}

// Rotated triangle record (TriRot, 16 floats = 64 bytes, leaf order): vertex v at floats
// [5v, 5v + 5) as (x, y, z, x, y), so that the components (kx, ky, kz) = (kz+1, kz+2, kz) mod 3
// of any ray are three consecutive floats starting at kx; float 15 = scene triangle index bits.
struct __attribute__((packed, aligned(4))) Packed3 {
    float x, y, z;
};

// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013, Sec. 4 (no
// backface culling; double-precision re-evaluation when an edge function is exactly zero), on a
// rotated triangle record `t`. Returns false when the ray misses the triangle's interior and
// edges; otherwise the hit distance is T / det (callers divide only for candidate hits).
__device__ __forceinline__ bool tri_watertight_td(const RayPre& r, const float* __restrict__ t, float* T_out,
                                                  float* det_out)
{
    // No fused multiply-adds here: an edge shared by two triangles must evaluate to exactly
    // opposite values in both, which fma(Cx, By, -Cy*Bx) vs fma(Bx, Cy, -By*Cx) does not give.
#pragma clang fp contract(off)
    const Packed3 a = *reinterpret_cast<const Packed3*>(t + r.rot);
    const Packed3 b = *reinterpret_cast<const Packed3*>(t + 5 + r.rot);
    const Packed3 c = *reinterpret_cast<const Packed3*>(t + 10 + r.rot);
    float Akx = a.x - r.o_rot.x, Aky = a.y - r.o_rot.y, Akz = a.z - r.o_rot.z;
    float Bkx = b.x - r.o_rot.x, Bky = b.y - r.o_rot.y, Bkz = b.z - r.o_rot.z;
    float Ckx = c.x - r.o_rot.x, Cky = c.y - r.o_rot.y, Ckz = c.z - r.o_rot.z;
    float Ax = Akx - r.Sx * Akz, Ay = Aky - r.Sy * Akz;
    float Bx = Bkx - r.Sx * Bkz, By = Bky - r.Sy * Bkz;
    float Cx = Ckx - r.Sx * Ckz, Cy = Cky - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    *T_out = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    *det_out = det;
    return true;
}

// Tests triangle record `t` (scene triangle `id`) against the current closest hit: updates hit and
// tfar; equal-distance ties go to the lower scene triangle index.
__device__ __forceinline__ void tri_hit(const RayPre& r, const float* __restrict__ t, float tnear, float& tfar,
                                        int32_t& hit)
{
    float T, det;
    if (!tri_watertight_td(r, t, &T, &det)) return;
    // divide only when an estimate of the distance is near (tnear, tfar]
    const float te = T * __builtin_amdgcn_rcpf(det);
    if (!(te >= tnear * 0.99999f && te <= tfar * 1.00001f)) return;
    const float th = T / det;
    const int32_t id = __float_as_int(t[15]);
    if (th > tnear && th <= tfar && (th < tfar || hit < 0 || id < hit)) {
        hit = id;
        tfar = th;
    }
}

__device__ __forceinline__ int32_t sel4(int32_t a0, int32_t a1, int32_t a2, int32_t a3, int k)
{
    return k == 0 ? a0 : (k == 1 ? a1 : (k == 2 ? a2 : a3));
}

// Closest (ANY = false) or any (ANY = true) hit in (tnear, tfar]. On a closest hit `tfar` is
// updated. Equal-distance ties resolve to the lower scene triangle index. `stack` points at this
// thread's first LDS stack slot; consecutive slots are `stride` ints apart.
// Returns the scene triangle index or -1.
template <bool ANY>
__device__ __forceinline__ int32_t trace(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                         const RayPre& r, float tnear, float& tfar, int32_t* stack,
                                         int stride)
{
    int32_t hit = -1;
    int32_t node = 0;
    int sp = 0;
    for (;;) {
        const float4* n = nodes + 4 * size_t(node);
        float4 a = n[0], b = n[1], c = n[2];
        int4 dd = *reinterpret_cast<const int4*>(n + 3);
        // slab tests as (lo - o) * inv: exact at a face through the origin, three roundings per
        // bound, so the far distance is widened by 1 + 2*gamma(3) (PBRT 3.9.2)
        float lx0 = (a.x - r.o.x) * r.inv.x, lx1 = (a.y - r.o.x) * r.inv.x;
        float ly0 = (a.z - r.o.y) * r.inv.y, ly1 = (a.w - r.o.y) * r.inv.y;
        float lz0 = (c.x - r.o.z) * r.inv.z, lz1 = (c.y - r.o.z) * r.inv.z;
        float rx0 = (b.x - r.o.x) * r.inv.x, rx1 = (b.y - r.o.x) * r.inv.x;
        float ry0 = (b.z - r.o.y) * r.inv.y, ry1 = (b.w - r.o.y) * r.inv.y;
        float rz0 = (c.z - r.o.z) * r.inv.z, rz1 = (c.w - r.o.z) * r.inv.z;
        float tl0 = fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)), fmaxf(fminf(lz0, lz1), tnear));
        float tl1 = fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)), fminf(fmaxf(lz0, lz1), tfar)) * 1.00000036f;
        float tr0 = fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)), fmaxf(fminf(rz0, rz1), tnear));
        float tr1 = fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)), fminf(fmaxf(rz0, rz1), tfar)) * 1.00000036f;
        bool hl = (tl0 <= tl1) && dd.z >= 0;
        bool hr = (tr0 <= tr1) && dd.w >= 0;
        // leaf children: test their triangles now
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            bool h = side == 0 ? hl : hr;
            int cnt = side == 0 ? dd.z : dd.w;
            if (h && cnt > 0) {
                int first = side == 0 ? dd.x : dd.y;
                for (int i = first; i < first + cnt; ++i) {
                    tri_hit(r, reinterpret_cast<const float*>(tris + 4 * size_t(i)), tnear, tfar, hit);
                    if (ANY && hit >= 0) return hit;
                }
                if (side == 0) hl = false;
                else hr = false;
            }
        }
        if (hl && hr) {
            bool left_first = tl0 <= tr0;
            stack[sp * stride] = left_first ? dd.y : dd.x;
            ++sp;
            node = left_first ? dd.x : dd.y;
        } else if (hl) {
            node = dd.x;
        } else if (hr) {
            node = dd.y;
        } else {
            if (sp == 0) break;
            --sp;
            node = stack[sp * stride];
        }
    }
    return hit;
}

// Traversal stack entry formats of the 4-wide BVH, kept in LDS, one column per thread:
//   kStack32   : 32-bit node index;
//   kStack16   : 16-bit node index (BVHs of at most 65,536 nodes): half the LDS, so more
//                wavefronts fit on a compute unit;
//   kStackKeyed: 16-bit node index | the upper 16 bits of the child's entry distance (a lower bound
//                of it, the float truncated to 7 mantissa bits): a closest-hit query skips a popped
//                child that starts beyond the current hit without fetching the node.
enum { kStack32 = 0, kStack16 = 1, kStackKeyed = 2 };

// Same query over the 4-wide BVH (Bvh4NodeHost: 8 float4 per node). Children that the ray enters
// are visited nearest first; the others are pushed far-to-near. Leaves are tested on the spot.
// `stack` is the block's LDS stack, `tid` this thread's column and `stride` the block size.
constexpr uint32_t kQEmptyDev = 7u;  // child reference count of an empty slot (Bvh4QNodeHost)

// Per-lane state of a 4-wide BVH query between node steps.
struct Trav4 {
    int32_t node;      // node to visit next (-1: none); while leaves are tested, the node they belong to
    int32_t sp;        // stack entries
    int32_t hit;       // closest (or any) hit so far, -1 if none
    float tfar;        // current far bound (the hit distance once there is a hit)
    uint32_t tri;      // leaf mode (tri < tri_end): next triangle of the current leaf
    uint32_t tri_end;  // end of the current leaf's triangles
    uint32_t leaves;   // leaf mode: entered leaf children of `node` not finished (the lowest bit is current)
    int32_t pnode;     // postponed leaves (closest hit): the node whose leaves `leaves` are
};

__device__ __forceinline__ void trav4_begin(Trav4& st, float tfar)
{
    st.node = 0;
    st.sp = 0;
    st.hit = -1;
    st.tfar = tfar;
    st.tri = 0;
    st.tri_end = 0;
    st.leaves = 0;
    st.pnode = -1;
}

// One step of a closest (ANY = false) or any (ANY = true) hit query over the quantized 4-wide BVH
// (Bvh4QNodeHost: 4 x 16 bytes per node). A step visits one node: the inner children the ray
// enters are pushed far-to-near (any hit: in slot order) and the nearest is visited next. A
// closest-hit query tests the node's entered leaves in the same step and then drops the inner
// children behind the hit. An any-hit query instead pushes every entered inner child and tests
// its leaves one triangle per step (leaf mode), popping the next node when they are done: lanes in
// leaf mode and lanes visiting nodes share each step, so a wavefront does not wait on a few lanes
// looping over their leaves (7 % faster shadow rays; for closest hits the lost culling costs
// more). Returns true when the query is complete (st.hit, st.tfar hold the answer). `stack` is the
// block's LDS stack, `tid` this thread's column and `stride` the block size.
template <bool ANY, int SK>
__device__ __forceinline__ bool trace4_step(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                            const RayPre& r, float tnear, Trav4& st, uint32_t* stack, int tid,
                                            int stride)
{
    static_assert(!(ANY && SK == kStackKeyed), "any-hit queries keep no entry distances");
    int32_t hit = st.hit;
    float tfar = st.tfar;
    int sp = st.sp;
    auto push = [&](int32_t child, float key) {
        if constexpr (SK == kStack32) {
            stack[sp * stride + tid] = uint32_t(child);
        } else if constexpr (SK == kStack16) {
            // entries 2k and 2k+1 of a thread share one 32-bit word: no two lanes write one word
            reinterpret_cast<uint16_t*>(stack)[((sp >> 1) * stride + tid) * 2 + (sp & 1)] = uint16_t(child);
        } else {
            stack[sp * stride + tid] = (__float_as_uint(key) & 0xFFFF0000u) | uint32_t(child);
        }
        ++sp;
    };
    // next node from the stack; false when the stack is empty (the query is complete). A keyed
    // entry whose box starts beyond the hit found since it was pushed is dropped (the same
    // 1 + 2^-21.5 margin as the box test, so no triangle at t <= tfar is skipped).
    auto pop = [&]() -> bool {
        while (sp > 0) {
            --sp;
            int32_t nd;
            if constexpr (SK == kStack32) {
                nd = int32_t(stack[sp * stride + tid]);
            } else if constexpr (SK == kStack16) {
                nd = reinterpret_cast<const uint16_t*>(stack)[((sp >> 1) * stride + tid) * 2 + (sp & 1)];
            } else {
                const uint32_t e = stack[sp * stride + tid];
                if (__uint_as_float(e & 0xFFFF0000u) > tfar * 1.00000036f) continue;
                nd = int32_t(e & 0xFFFFu);
            }
            st.node = nd;
            return true;
        }
        return false;
    };
    bool done = false;
    if (ANY && st.tri < st.tri_end) {
        // leaf mode (any-hit queries): one triangle
        tri_hit(r, reinterpret_cast<const float*>(tris + 4 * size_t(st.tri)), tnear, tfar, hit);
        if (ANY && hit >= 0) {
            st.hit = hit;
            st.tfar = tfar;
            return true;
        }
        if (++st.tri == st.tri_end) {
            const uint32_t rest = st.leaves & (st.leaves - 1u);
            st.leaves = rest;
            if (rest) {
                // the next entered leaf of this node: its reference is re-read from the node
                const uint4 w3 = reinterpret_cast<const uint4*>(nodes)[4 * size_t(st.node) + 3];
                const int k = __builtin_ctz(rest);
                const uint32_t ref = uint32_t(sel4(int32_t(w3.x), int32_t(w3.y), int32_t(w3.z), int32_t(w3.w), k));
                st.tri = ref & 0x1FFFFFFFu;
                st.tri_end = st.tri + (ref >> 29);
            } else {
                done = !pop();
            }
        }
        st.sp = sp;
        st.hit = hit;
        st.tfar = tfar;
        return done;
    }
    // quantized node (Bvh4QNodeHost): origin, grid steps, child box bytes, child references
    const uint4* n = reinterpret_cast<const uint4*>(nodes) + 4 * size_t(st.node);
    const uint4 w0 = n[0], w1 = n[1], w2 = n[2], w3 = n[3];
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const float sx = __uint_as_float(w2.z), sy = __uint_as_float(w2.w), sz = __uint_as_float(w0.w);
    const uint32_t ref[4] = {w3.x, w3.y, w3.z, w3.w};
    int cha[4], cna[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        cha[k] = int(ref[k] & 0x1FFFFFFFu);
        const int c = int(ref[k] >> 29);
        cna[k] = c == int(kQEmptyDev) ? -1 : c;
    }
    // Slab distances in ray space: (origin + q * step - o) * inv = q * a + b with a = step * inv
    // (exact: step is a power of two) and b = origin * inv - o * inv, so each plane costs one fused
    // multiply-add. The entry plane of an axis (the lower one for inv >= 0) is evaluated with
    // b - k |inv| and the exit plane with b + k |inv|, a margin that bounds every rounding against
    // the plane as the host rounds it (ray_pre). Entry and exit need no min / max.
    const float ax = sx * r.inv.x, ay = sy * r.inv.y, az = sz * r.inv.z;
    const float bx = fmaf(ox, r.inv.x, r.c.x), by = fmaf(oy, r.inv.y, r.c.y), bz = fmaf(oz, r.inv.z, r.c.z);
    const float bx0 = fmaf(-r.k, fabsf(r.inv.x), bx), bx1 = fmaf(r.k, fabsf(r.inv.x), bx);
    const float by0 = fmaf(-r.k, fabsf(r.inv.y), by), by1 = fmaf(r.k, fabsf(r.inv.y), by);
    const float bz0 = fmaf(-r.k, fabsf(r.inv.z), bz), bz1 = fmaf(r.k, fabsf(r.inv.z), bz);
    const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
    const uint32_t nx = px ? w1.x : w1.y, fx = px ? w1.y : w1.x;  // entry / exit plane bytes
    const uint32_t ny = py ? w1.z : w1.w, fy = py ? w1.w : w1.z;
    const uint32_t nz = pz ? w2.x : w2.y, fz = pz ? w2.y : w2.x;
    float key[4];
    int idx[4];
    uint32_t leaves = 0;  // children whose box the ray enters and that are leaves
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float x0 = fmaf(float((nx >> (8 * k)) & 0xFFu), ax, bx0), x1 = fmaf(float((fx >> (8 * k)) & 0xFFu), ax, bx1);
        const float y0 = fmaf(float((ny >> (8 * k)) & 0xFFu), ay, by0), y1 = fmaf(float((fy >> (8 * k)) & 0xFFu), ay, by1);
        const float z0 = fmaf(float((nz >> (8 * k)) & 0xFFu), az, bz0), z1 = fmaf(float((fz >> (8 * k)) & 0xFFu), az, bz1);
        const float t0 = fmaxf(fmaxf(x0, y0), fmaxf(z0, tnear));
        const float t1 = fminf(fminf(x1, y1), fminf(z1, tfar)) * 1.00000036f;
        bool h = t0 <= t1 && cna[k] >= 0;
        if (h && cna[k] > 0) leaves |= 1u << k;
        h = h && cna[k] == 0;
        key[k] = h ? t0 : __int_as_float(0x7f800000);
        idx[k] = h ? cha[k] : -1;
    }
    if (!ANY && leaves) {
        // closest hit: the triangles of every entered leaf now, one per iteration (a wavefront runs
        // as many iterations as its busiest lane needs), so that the inner children behind a hit
        // found here are not entered at all
        int k = __builtin_ctz(leaves);
        int i = sel4(cha[0], cha[1], cha[2], cha[3], k);
        int e = i + sel4(cna[0], cna[1], cna[2], cna[3], k);
        for (;;) {
            tri_hit(r, reinterpret_cast<const float*>(tris + 4 * size_t(i)), tnear, tfar, hit);
            if (++i == e) {
                leaves &= leaves - 1u;
                if (!leaves) break;
                k = __builtin_ctz(leaves);
                i = sel4(cha[0], cha[1], cha[2], cha[3], k);
                e = i + sel4(cna[0], cna[1], cna[2], cna[3], k);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (!(key[c] <= tfar * 1.00000036f)) {
                key[c] = __int_as_float(0x7f800000);
                idx[c] = -1;
            }
    }
    if constexpr (!ANY) {
        // sort (key, idx) ascending: 5-comparator network; nearest child first
#define TR_CSWAP(a, b)                                                            \
    if (key[b] < key[a]) {                                                        \
        float tk = key[a]; key[a] = key[b]; key[b] = tk;                          \
        int ti = idx[a]; idx[a] = idx[b]; idx[b] = ti;                            \
    }
        TR_CSWAP(0, 1) TR_CSWAP(2, 3) TR_CSWAP(0, 2) TR_CSWAP(1, 3) TR_CSWAP(1, 2)
#undef TR_CSWAP
    } else {
        // any hit: order does not matter; move an entered child to slot 0
#pragma unroll
        for (int c = 1; c < 4; ++c)
            if (idx[0] < 0 && idx[c] >= 0) {
                idx[0] = idx[c];
                idx[c] = -1;
            }
    }
    // push the entered inner children (far first); descend into the nearest, or, for an any-hit
    // query with entered leaves, push every one of them and test the leaves from the next step on
#pragma unroll
    for (int k = 3; k >= 1; --k)
        if (idx[k] >= 0) push(idx[k], key[k]);
    if (ANY && leaves) {
        if (idx[0] >= 0) push(idx[0], key[0]);
        const int k = __builtin_ctz(leaves);
        st.leaves = leaves;
        st.tri = uint32_t(sel4(cha[0], cha[1], cha[2], cha[3], k));
        st.tri_end = st.tri + uint32_t(sel4(cna[0], cna[1], cna[2], cna[3], k));
    } else if (idx[0] >= 0) {
        st.node = idx[0];
    } else {
        done = !pop();
    }
    st.sp = sp;
    st.hit = hit;
    st.tfar = tfar;
    return done;
}
// ---- closest hit of a wavefront's rays as one packet ------------------------------------------
// The camera rays of one pixel's samples nearly share their traversal. Here the wavefront visits one
// node at a time for all its lanes: a node is visited when the ray of any active lane enters it, so
// the node index and the traversal stack are uniform, node and triangle reads are the same address
// for every lane, and every lane runs the same instruction stream (a per-lane query runs each node
// step with about half the lanes of a wavefront active). Each lane keeps its own closest hit and far
// bound, tests the triangles of a leaf its ray enters, and drops a child whose box starts beyond its
// hit. A lane's answer is the one trace4_step gives: the closest hit does not depend on the order of
// the visits or on visits of boxes the ray misses (equal distances resolve to the lower triangle
// index). Children are visited nearest first by the entry distances of the first active lane.
// `col` is the stack column of the wavefront's first lane (all lanes write the same entries).
template <int SK>
__device__ __forceinline__ void trace4_packet(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                              const RayPre& r, float tnear, bool active, float& tfar, int32_t& hit,
                                              uint32_t* stack, int col, int stride)
{
    const uint64_t amask = __ballot(active);
    if (amask == 0ull) return;
    const int lead = int(__builtin_ctzll(amask));
    auto push = [&](int sp, int32_t child) {
        if constexpr (SK == kStack16)
            reinterpret_cast<uint16_t*>(stack)[((sp >> 1) * stride + col) * 2 + (sp & 1)] = uint16_t(child);
        else
            stack[sp * stride + col] = uint32_t(child);
    };
    auto pop = [&](int sp) -> int32_t {
        if constexpr (SK == kStack16)
            return int32_t(reinterpret_cast<const uint16_t*>(stack)[((sp >> 1) * stride + col) * 2 + (sp & 1)]);
        else
            return int32_t(stack[sp * stride + col] & (SK == kStackKeyed ? 0xFFFFu : 0xFFFFFFFFu));
    };
    // the node is the same for every lane: the device pass reads it through the constant address
    // space, so that the compiler uses scalar loads (one per wavefront, into scalar registers); the
    // host pass, which only parses this function, has no such address space
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const uint4 cuint4;
#else
    typedef const uint4 cuint4;
#endif
    int32_t node = 0;
    int sp = 0;
    for (;;) {
        node = __builtin_amdgcn_readfirstlane(node);
        const cuint4* n = (const cuint4*)(reinterpret_cast<const uint4*>(nodes) + 4 * size_t(node));
        const uint4 w0 = n[0], w1 = n[1], w2 = n[2], w3 = n[3];
        const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
        const float sx = __uint_as_float(w2.z), sy = __uint_as_float(w2.w), sz = __uint_as_float(w0.w);
        const uint32_t ref[4] = {w3.x, w3.y, w3.z, w3.w};
        int cha[4], cna[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cha[k] = int(ref[k] & 0x1FFFFFFFu);
            const int c = int(ref[k] >> 29);
            cna[k] = c == int(kQEmptyDev) ? -1 : c;
        }
        // the slab test of trace4_step (same planes, same margins)
        const float ax = sx * r.inv.x, ay = sy * r.inv.y, az = sz * r.inv.z;
        const float bx = fmaf(ox, r.inv.x, r.c.x), by = fmaf(oy, r.inv.y, r.c.y), bz = fmaf(oz, r.inv.z, r.c.z);
        const float bx0 = fmaf(-r.k, fabsf(r.inv.x), bx), bx1 = fmaf(r.k, fabsf(r.inv.x), bx);
        const float by0 = fmaf(-r.k, fabsf(r.inv.y), by), by1 = fmaf(r.k, fabsf(r.inv.y), by);
        const float bz0 = fmaf(-r.k, fabsf(r.inv.z), bz), bz1 = fmaf(r.k, fabsf(r.inv.z), bz);
        const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
        const uint32_t nx = px ? w1.x : w1.y, fx = px ? w1.y : w1.x;
        const uint32_t ny = py ? w1.z : w1.w, fy = py ? w1.w : w1.z;
        const uint32_t nz = pz ? w2.x : w2.y, fz = pz ? w2.y : w2.x;
        float key[4];
        bool ent[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float x0 = fmaf(float((nx >> (8 * k)) & 0xFFu), ax, bx0), x1 = fmaf(float((fx >> (8 * k)) & 0xFFu), ax, bx1);
            const float y0 = fmaf(float((ny >> (8 * k)) & 0xFFu), ay, by0), y1 = fmaf(float((fy >> (8 * k)) & 0xFFu), ay, by1);
            const float z0 = fmaf(float((nz >> (8 * k)) & 0xFFu), az, bz0), z1 = fmaf(float((fz >> (8 * k)) & 0xFFu), az, bz1);
            const float t0 = fmaxf(fmaxf(x0, y0), fmaxf(z0, tnear));
            const float t1 = fminf(fminf(x1, y1), fminf(z1, tfar)) * 1.00000036f;
            ent[k] = active && t0 <= t1 && cna[k] >= 0;
            key[k] = t0;
        }
        // leaves first: the wavefront loops over the triangles of every leaf some lane entered;
        // each lane tests those of the leaves its ray entered
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (cna[k] > 0 && __ballot(ent[k]) != 0ull) {
                const int i0 = __builtin_amdgcn_readfirstlane(cha[k]), i1 = i0 + __builtin_amdgcn_readfirstlane(cna[k]);
                for (int i = i0; i < i1; ++i)
                    if (ent[k]) tri_hit(r, reinterpret_cast<const float*>(tris + 4 * size_t(i)), tnear, tfar, hit);
            }
        }
        // inner children that some lane enters before its hit, ordered by the first active lane's
        // entry distances (children that lane does not enter last)
        float uk[4];
        int ui[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool e = ent[k] && cna[k] == 0 && key[k] <= tfar * 1.00000036f;
            const bool any = __ballot(e) != 0ull;
            const float lk = __shfl(e ? key[k] : __int_as_float(0x7f800000), lead);
            uk[k] = any ? lk : __int_as_float(0x7f800000);
            ui[k] = any ? __builtin_amdgcn_readfirstlane(cha[k]) : -1;
        }
#define TR_PSWAP(a, b)                                                                \
    if (ui[b] >= 0 && (ui[a] < 0 || uk[b] < uk[a])) {                                 \
        float tk = uk[a]; uk[a] = uk[b]; uk[b] = tk;                                  \
        int ti = ui[a]; ui[a] = ui[b]; ui[b] = ti;                                    \
    }
        TR_PSWAP(0, 1) TR_PSWAP(2, 3) TR_PSWAP(0, 2) TR_PSWAP(1, 3) TR_PSWAP(1, 2)
#undef TR_PSWAP
        if (ui[0] < 0) {
            if (sp == 0) break;
            --sp;
            node = pop(sp);
            continue;
        }
#pragma unroll
        for (int k = 3; k >= 1; --k)
            if (ui[k] >= 0) {
                push(sp, ui[k]);
                ++sp;
            }
        node = ui[0];
    }
}

// ---- closest hit with postponed leaves -------------------------------------------------------
// A closest-hit query whose wavefront runs node steps and triangle steps as separate, wavefront-wide
// phases: a node step records the node's entered leaves as the lane's pending leaves instead of
// testing them, and the lane goes on to the nearest inner child (or pops the keyed stack) while
// its leaves wait; the caller runs a triangle phase (one triangle per lane with pending leaves)
// once enough lanes hold some, or when no lane can take a node step. A lane that meets a second
// node with entered leaves while it still holds pending ones waits on that node and visits it
// again after its leaves are done (with the smaller far bound). Nodes and triangles then run with
// most lanes active instead of every node step carrying the triangle loop of a few lanes.
// Needs the keyed stack (kStackKeyed): entries behind a hit found later are dropped on pop.

// true when the lane can take a node step now
__device__ __forceinline__ bool trav4_pp_can_step(const Trav4& st, bool blocked) { return st.node >= 0 && !blocked; }

// One node step; `blocked` is set when the node has entered leaves while the lane holds pending
// ones (the state is left unchanged: the node is visited again once they are tested). Returns true
// when the query is complete.
template <int SK>
__device__ __forceinline__ bool trace4_node_pp(const float4* __restrict__ nodes, const RayPre& r, float tnear, Trav4& st,
                                               uint32_t* stack, int tid, int stride, bool& blocked)
{
    static_assert(SK == kStackKeyed || SK == kStack32, "postponed leaves use the keyed (or 32-bit) stack");
    const float tfar = st.tfar;
    int sp = st.sp;
    const uint4* n = reinterpret_cast<const uint4*>(nodes) + 4 * size_t(st.node);
    const uint4 w0 = n[0], w1 = n[1], w2 = n[2], w3 = n[3];
    const float ox = __uint_as_float(w0.x), oy = __uint_as_float(w0.y), oz = __uint_as_float(w0.z);
    const float sx = __uint_as_float(w2.z), sy = __uint_as_float(w2.w), sz = __uint_as_float(w0.w);
    const uint32_t ref[4] = {w3.x, w3.y, w3.z, w3.w};
    int cha[4], cna[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        cha[k] = int(ref[k] & 0x1FFFFFFFu);
        const int c = int(ref[k] >> 29);
        cna[k] = c == int(kQEmptyDev) ? -1 : c;
    }
    // the slab test of trace4_step (same planes, same margins)
    const float ax = sx * r.inv.x, ay = sy * r.inv.y, az = sz * r.inv.z;
    const float bx = fmaf(ox, r.inv.x, r.c.x), by = fmaf(oy, r.inv.y, r.c.y), bz = fmaf(oz, r.inv.z, r.c.z);
    const float bx0 = fmaf(-r.k, fabsf(r.inv.x), bx), bx1 = fmaf(r.k, fabsf(r.inv.x), bx);
    const float by0 = fmaf(-r.k, fabsf(r.inv.y), by), by1 = fmaf(r.k, fabsf(r.inv.y), by);
    const float bz0 = fmaf(-r.k, fabsf(r.inv.z), bz), bz1 = fmaf(r.k, fabsf(r.inv.z), bz);
    const bool px = r.inv.x >= 0.0f, py = r.inv.y >= 0.0f, pz = r.inv.z >= 0.0f;
    const uint32_t nx = px ? w1.x : w1.y, fx = px ? w1.y : w1.x;
    const uint32_t ny = py ? w1.z : w1.w, fy = py ? w1.w : w1.z;
    const uint32_t nz = pz ? w2.x : w2.y, fz = pz ? w2.y : w2.x;
    float key[4];
    int idx[4];
    uint32_t leaves = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float x0 = fmaf(float((nx >> (8 * k)) & 0xFFu), ax, bx0), x1 = fmaf(float((fx >> (8 * k)) & 0xFFu), ax, bx1);
        const float y0 = fmaf(float((ny >> (8 * k)) & 0xFFu), ay, by0), y1 = fmaf(float((fy >> (8 * k)) & 0xFFu), ay, by1);
        const float z0 = fmaf(float((nz >> (8 * k)) & 0xFFu), az, bz0), z1 = fmaf(float((fz >> (8 * k)) & 0xFFu), az, bz1);
        const float t0 = fmaxf(fmaxf(x0, y0), fmaxf(z0, tnear));
        const float t1 = fminf(fminf(x1, y1), fminf(z1, tfar)) * 1.00000036f;
        bool h = t0 <= t1 && cna[k] >= 0;
        if (h && cna[k] > 0) leaves |= 1u << k;
        h = h && cna[k] == 0;
        key[k] = h ? t0 : __int_as_float(0x7f800000);
        idx[k] = h ? cha[k] : -1;
    }
    if (leaves && st.leaves) {
        blocked = true;  // test the pending leaves first, then visit this node again
        return false;
    }
    if (leaves) {
        const int k = __builtin_ctz(leaves);
        st.pnode = st.node;
        st.leaves = leaves;
        st.tri = uint32_t(sel4(cha[0], cha[1], cha[2], cha[3], k));
        st.tri_end = st.tri + uint32_t(sel4(cna[0], cna[1], cna[2], cna[3], k));
    }
#define TR_CSWAP(a, b)                                                            \
    if (key[b] < key[a]) {                                                        \
        float tk = key[a]; key[a] = key[b]; key[b] = tk;                          \
        int ti = idx[a]; idx[a] = idx[b]; idx[b] = ti;                            \
    }
    TR_CSWAP(0, 1) TR_CSWAP(2, 3) TR_CSWAP(0, 2) TR_CSWAP(1, 3) TR_CSWAP(1, 2)
#undef TR_CSWAP
#pragma unroll
    for (int k = 3; k >= 1; --k)
        if (idx[k] >= 0) {
            if constexpr (SK == kStack32) stack[sp * stride + tid] = uint32_t(idx[k]);
            else stack[sp * stride + tid] = (__float_as_uint(key[k]) & 0xFFFF0000u) | uint32_t(idx[k]);
            ++sp;
        }
    int32_t next = idx[0];
    while (next < 0 && sp > 0) {
        --sp;
        const uint32_t e = stack[sp * stride + tid];
        if constexpr (SK == kStack32) {
            next = int32_t(e);
        } else {
            if (__uint_as_float(e & 0xFFFF0000u) > tfar * 1.00000036f) continue;
            next = int32_t(e & 0xFFFFu);
        }
    }
    st.node = next;
    st.sp = sp;
    return next < 0 && st.leaves == 0;
}

// One triangle of the lane's pending leaves. Returns true when the query is complete.
__device__ __forceinline__ bool trace4_leaf_pp(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                               const RayPre& r, float tnear, Trav4& st, bool& blocked)
{
    tri_hit(r, reinterpret_cast<const float*>(tris + 4 * size_t(st.tri)), tnear, st.tfar, st.hit);
    if (++st.tri == st.tri_end) {
        const uint32_t rest = st.leaves & (st.leaves - 1u);
        st.leaves = rest;
        if (rest) {
            const uint4 w3 = reinterpret_cast<const uint4*>(nodes)[4 * size_t(st.pnode) + 3];
            const int k = __builtin_ctz(rest);
            const uint32_t ref = uint32_t(sel4(int32_t(w3.x), int32_t(w3.y), int32_t(w3.z), int32_t(w3.w), k));
            st.tri = ref & 0x1FFFFFFFu;
            st.tri_end = st.tri + (ref >> 29);
        } else {
            blocked = false;
        }
    }
    return st.leaves == 0 && st.node < 0;
}

template <bool ANY, int SK>
__device__ __forceinline__ int32_t trace4(const float4* __restrict__ nodes, const float4* __restrict__ tris,
                                          const RayPre& r, float tnear, float& tfar, uint32_t* stack, int tid,
                                          int stride)
{
    Trav4 st;
    trav4_begin(st, tfar);
    while (!trace4_step<ANY, SK>(nodes, tris, r, tnear, st, stack, tid, stride)) {
    }
    tfar = st.tfar;
    return st.hit;
}

}  // namespace trd
