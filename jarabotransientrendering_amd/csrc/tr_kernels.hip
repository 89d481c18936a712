// Wavefront transient path tracer for gfx950 (MI355X) and the C ABI declared in include/tr_api.h.
//
// Per wavefront of `wave` paths:
//   camera_kernel                 camera ray + per-path PCG32 streams       (TransientRenderer.h:130-163,
//                                                                            StratifiedSampler.h:47-89,
//                                                                            PinholeCamera.h:82-108)
//   for each bounce:
//     extend_kernel               closest hit of every queued ray           (World::first_intersection)
//     shade_kernel                free-flight correction, vertex, NEE light
//                                 sample -> shadow queue, roulette, next
//                                 direction -> next ray queue               (BidirectionalPathTracing.h:486-692)
//     shadow_kernel               occlusion of every shadow ray             (World::intersects)
//     film_kernel                 visible contributions binned into the
//                                 film's compensated FP32 (sum, comp) pairs (StreakCameraFilm.h:139-192)
//
// Path state lives in HBM as structure-of-arrays float4/uint4 records in two buffers: the rays of
// bounce k are records [0, count) of buffer k & 1, and shade writes the paths that continue,
// compacted in queue order, into the other buffer, so every kernel reads and writes its records
// densely. Queue counters live in device memory so no host round trip is needed between kernels.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "tr_api.h"
#include "tr_device.h"
#include "tr_guide.h"
#include "tr_guide_host.h"
#include "tr_internal.h"

using namespace trd;

namespace {

constexpr int kTraceBlock = 128;   // traversal kernels: 128 threads, STACK ints of LDS each
constexpr int kBlock = 256;        // other kernels
constexpr int kShadeBlock = 1024;  // shade: one queue atomic per 1024 paths
constexpr uint32_t kDefaultWave = 1u << 25;
constexpr uint32_t kFetchMax = 512;  // rays per work fetch of a traversal wavefront (upper bound)
constexpr uint32_t kTraceRaysPerThread = 8;   // lower bound on a full wavefront's rays per traversal thread
constexpr uint32_t kTraceMinPerCu = 4;        // traversal blocks per CU below which that bound yields

// Rays per work fetch: large enough that same-address fetch atomics stay rare, small enough that
// every wavefront of the grid still gets several fetches when the queue is short.
__device__ __forceinline__ uint32_t fetch_chunk(uint32_t count)
{
    uint32_t waves = gridDim.x * (blockDim.x >> 6);
    uint32_t c = count / (waves * 4u);
    c = c < 64u ? 64u : (c > kFetchMax ? kFetchMax : c);
    return c & ~63u;
}
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kInv4Pi = 0.0795774715459476678844f;  // .25 * M_1_PI (IsotropicPhaseFunction.h:43-47)
constexpr float kSigmaVisibility = 1e-5f;             // _SIGMA_VISIBILITY_ (bunnykiller.h:76-78)
// material bits of a scene triangle: index | two-sided material | smooth (per-corner normals)
constexpr int32_t kMatTwoSided = 1 << 30;
constexpr int32_t kMatSmooth = 1 << 29;
constexpr int32_t kMatIndex = kMatSmooth - 1;
// multibounce films: bits of a shadow record's tag that hold its layer (at most 32 layers)
constexpr uint32_t kMaxLayersTagBits = kMbLayerTagBits;

// counters[] layout (uint32 words): queue lengths and per-launch statistics in the first 256 bytes,
// then kParts partitions of 256 bytes each holding the work-fetch cursors of the persistent
// traversal grids and the visible-contribution count. Atomics to one address are serialised, and
// the wavefronts of a launch fetch work all the time: blocks are dealt to the 8 XCDs in turn, so
// partition blockIdx.x % kParts takes the fetches of one XCD's blocks on its own 256-byte line.
constexpr uint32_t kParts = 8;
constexpr uint32_t kPartBase = 64;
constexpr uint32_t kPartWords = 64;
// Queue lengths come in two 64-bit pairs, one per record buffer k: (rays of buffer k, shadow
// rays made from the paths of the other buffer), so that shade appends to both of its output
// queues with one atomic.
enum {
    C_Q0 = 0,                                            // pairs: words 2k (rays), 2k + 1 (shadow rays)
    C_STAT_EXT = 4, C_STAT_SHADOW = 6,                   // 64-bit statistics
    C_WORDS = kPartBase + kParts * kPartWords
};
__host__ __device__ constexpr uint32_t qword(int k) { return uint32_t(2 * k); }
__host__ __device__ constexpr uint32_t shword(int k) { return uint32_t(2 * k + 1); }
enum { P_FETCH_EXT = 0, P_FETCH_SHADOW = 1, P_STAT_CONTRIB = 2 };  // words of a partition (P_STAT_CONTRIB: 64-bit)
__host__ __device__ constexpr uint32_t part_word(uint32_t part, uint32_t w) { return kPartBase + part * kPartWords + w; }

struct SceneDev {
    const float4* nodes;       // binary BVH (4 float4 per node)
    const float4* nodes4;      // quantized 4-wide BVH (4 x 16 bytes per node, Bvh4QNodeHost)
    const float4* tris;        // rotated triangle records in leaf order (4 float4 each, see TriRot)
    const float4* tri_normal;  // per scene triangle: geometric normal xyz, w = material bits (int)
    const float4* tri_vn;      // smooth triangles: p0, e1, e2, n0, n1, n2 (null if none)
    const float4* mats;        // Kd xyz, w = mean albedo used by the roulette
    const LightHost* lights;
    int32_t n_lights;
    int32_t has_medium;
    float3 sigma_s, sigma_t;
    float lambda;              // mean extinction (HomogeneousMedium.h:50-51)
    float medium_albedo;       // mean of sigma_s / sigma_t
    float plane_bound;         // largest |plane| of the quantized BVH (qbvh_plane_bound)
};

struct RenderDev {
    // film
    float* film;               // compensated sums [rows][W][T][3]: sum ...
    float* film_comp;          // ... and the rounding errors of its additions (film_add3)
    int32_t W, H, T, row0, spp, spp2, n_inc;
    float exposure, offset, total_time, inv_inc;
    int32_t unwarp, max_bounces;
    // camera
    float3 cpos, nv, uv, vv;
    float fd, xw2, yw2, ratio, inv_spp;
    uint64_t seed;
    // MultibounceStreakFilm: `layers` bounce films of layer_stride bins each (0 = plain film)
    int32_t layers;
    uint32_t layer_stride;
    float* steady;             // [rows][W][3] every sample, any time (multibounce films) ...
    float* steady_comp;        // ... and its compensation
    float* layer_w;            // [layers][rows][W] camera samples with samples of each layer
    uint32_t* cs_bits;         // per camera sample of the wavefront: layers that received samples
    // Transient sampling mode (-time-sampling): free-flight rate 4 / (T * exposure); 0 = steady state
    float ts_lambda;
    int32_t time_sampling;
};

struct WaveDev {
    // path records of the current ([cur]) and the next ([cur ^ 1]) bounce. The camera kernel writes
    // no origin and no throughput: every camera ray starts at the camera with unit throughput at
    // time 0, so the first bounce takes them from the camera set-up (and the sampled free-flight
    // distance from extra.w, which holds the unwarp distance only from the first hit on).
    float4* ray_o[2]; // origin xyz, w = sampled free-flight distance of this segment (inf without medium)
    float4* ray_d[2]; // direction xyz, w = batched renders: job (uint bits); otherwise the path's
                      // index in its wavefront (uint bits)
    float4* thr[2];   // throughput xyz (with the pending BSDF/phase weight and roulette), w = time
    uint4* misc[2];   // x = path id lo, y = path id hi, z/w = path-stream PCG32 state
    uint4* extra[2];  // x/y = light-stream PCG32 state, z = level | film pixel << 8, w = unwarp distance (float bits)
    float2* hit;      // per record of the current bounce: x = distance, y = scene triangle (int bits) or -1
    float4* sh_o;     // light point xyz, w = distance
    float4* sh_d;     // direction xyz, w = film element index (uint bits)
    float4* sh_c;     // contribution rgb, w = film pixel (uint bits)
    uint8_t* sh_vis;  // 1 if the shadow ray is unoccluded
    uint32_t* sh_tag; // guided renders: training record of the contribution (or ~0)
    uint32_t* sh_job; // batched renders: job of each shadow record
    struct WaveCtl* ctl;
    uint32_t* counters;
    float3 cpos;      // single-job renders: camera position (batched renders: their job's R.cpos)
};

// One render job of a batched launch: its scene, film/camera set-up and first global path index.
struct JobDev {
    SceneDev S;
    RenderDev R;
    uint64_t path_begin;
};

// Paths of the current wavefront: [base, base + n) of the launch's global path range. Kept in
// device memory so that the bounce loop of a wavefront can be replayed from a captured graph.
struct WaveCtl {
    uint64_t base;
    uint32_t n;
    uint32_t pad;
};

__global__ void set_wave_kernel(WaveCtl* ctl, uint64_t base, uint32_t n)
{
    ctl->base = base;
    ctl->n = n;
    ctl->pad = 0u;
}

__device__ __forceinline__ uint64_t u64(uint32_t lo, uint32_t hi) { return uint64_t(lo) | (uint64_t(hi) << 32); }

// Free-flight distance -log(eps) / lambda (BidirectionalPathTracing.h:258-259, float log): the
// log is taken in double and rounded, which gives the correctly rounded float log that host libms
// return, so that a path's distances match the host's bit for bit in all but rare cases.
__device__ __forceinline__ float free_flight(float eps, float lambda)
{
    return -float(log(double(eps))) / lambda;
}

// o + d * t rounded as two operations (no fused multiply-add), as the reference computes vertex
// positions (Ray::get_position): the time-angular sampler's geometry is ill-conditioned near
// backward scattering, so positions must match the host's to the last bit there.
__device__ __forceinline__ f3 point_on_ray(f3 o, f3 d, float t)
{
#pragma clang fp contract(off)
    return mk(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
}

// ---------------------------------------------------------------------------------------------
// camera: one thread per path of the wavefront
template <bool BATCH>
__global__ void __launch_bounds__(kBlock) camera_kernel(SceneDev S0, RenderDev R0, const JobDev* __restrict__ jobs,
                                                        uint32_t n_jobs, WaveDev Wv)
{
    const WaveCtl ctl = *Wv.ctl;
    const uint32_t n = ctl.n;
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot == 0) Wv.counters[C_Q0] = n;
    if (slot < kParts) Wv.counters[part_word(slot, P_FETCH_EXT)] = 0;
    if (slot >= n) return;
    // a job holds fewer than 2^32 paths: 32-bit index arithmetic within the job
    uint32_t jid = 0, pid_local;
    if constexpr (BATCH) {
        // job of this path: the last job whose first path is <= the global path index
        const uint64_t gidx = ctl.base + slot;
        uint32_t lo = 0, hi = n_jobs;
        while (hi - lo > 1u) {
            uint32_t mid = (lo + hi) >> 1;
            if (jobs[mid].path_begin <= gidx) lo = mid;
            else hi = mid;
        }
        jid = lo;
        pid_local = uint32_t(gidx - jobs[lo].path_begin);
    } else {
        pid_local = uint32_t(ctl.base) + slot;
    }
    const SceneDev& S = BATCH ? jobs[jid].S : S0;
    const RenderDev& R = BATCH ? jobs[jid].R : R0;
    uint32_t cs_local = pid_local / uint32_t(R.n_inc);
    uint32_t inc_i = pid_local - cs_local * uint32_t(R.n_inc);
    uint32_t pix_local = cs_local / uint32_t(R.spp2);   // film pixel (row relative)
    uint32_t sub = cs_local - pix_local * uint32_t(R.spp2);
    uint32_t rowl = pix_local / uint32_t(R.W);
    uint32_t xr = pix_local - rowl * uint32_t(R.W);
    uint32_t y = rowl + uint32_t(R.row0);
    uint32_t sy = sub / uint32_t(R.spp);
    uint32_t sx = sub - sy * uint32_t(R.spp);
    uint64_t cs = (uint64_t(y) * uint64_t(R.W) + xr) * uint64_t(R.spp2) + sub;
    uint64_t pid = cs * uint64_t(R.n_inc) + inc_i;

    // StratifiedSampler::get_next_sample: jittered position clamped inside the pixel
    uint64_t cinc = pcg_inc((1ull << 62) | cs);
    uint64_t cst = pcg_seed(R.seed, (1ull << 62) | cs);
    float jx = pcg_float(cst, cinc), jy = pcg_float(cst, cinc);
    float px = fminf(float(xr) + (float(sx) + jx) * R.inv_spp, __uint_as_float(__float_as_uint(float(xr + 1)) - 1u));
    float py = fminf(float(y) + (float(sy) + jy) * R.inv_spp, __uint_as_float(__float_as_uint(float(y + 1)) - 1u));
    // Film::window_coords2image_coords + PinholeCamera::get_ray_direction
    float ic0 = (px - R.xw2) / float(R.W);
    float ic1 = (py - R.yw2) / float(R.H) / R.ratio;
    f3 nv = mk(R.nv.x, R.nv.y, R.nv.z), uv = mk(R.uv.x, R.uv.y, R.uv.z), vv = mk(R.vv.x, R.vv.y, R.vv.z);
    f3 dir = nv * R.fd + uv * (2.0f * ic0) - vv * (2.0f * ic1);
    dir = normalize(normalize(dir));  // get_ray_direction normalises, Ray3D normalises again

    uint64_t pinc = pcg_inc(2 * pid);
    uint64_t pst = pcg_seed(R.seed, 2 * pid);
    uint64_t nst = pcg_seed(R.seed, 2 * pid + 1);
    float sampled = __int_as_float(0x7f800000);
    if (S.has_medium) {
        sampled = free_flight(pcg_float(pst, pinc), R.time_sampling ? R.ts_lambda : S.lambda);
    }
    Wv.ray_d[0][slot] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(BATCH ? jid : slot));
    Wv.misc[0][slot] = make_uint4(uint32_t(pid), uint32_t(pid >> 32), uint32_t(pst), uint32_t(pst >> 32));
    Wv.extra[0][slot] = make_uint4(uint32_t(nst), uint32_t(nst >> 32), pix_local << 8, __float_as_uint(sampled));
}

// ---------------------------------------------------------------------------------------------
// Traversal variants: the binary BVH, or the 4-wide BVH with 32-bit, 16-bit or keyed 16-bit stack
// entries (trd::kStack*). Any-hit queries of a keyed launch use the 16-bit format.
// kTravWidePP: keyed stack and closest hits with postponed leaves (trace4_node_pp / trace4_leaf_pp);
// its any-hit launches use the 16-bit stack.
enum { kTravBinary = 0, kTravWide = 1, kTravWide16 = 2, kTravWideKeyed = 3, kTravWidePP = 4 };
// postponed leaves: a triangle phase runs once this many lanes of a wavefront hold pending leaves
constexpr int kLeafPhaseMin = 32;

// LDS words per thread of a traversal block for a stack of STACK entries
template <int TRAV, int STACK>
constexpr int stack_words() { return TRAV == kTravWide16 ? (STACK + 1) / 2 : STACK; }

template <bool ANY, int TRAV>
__device__ __forceinline__ int32_t trace_scene(const SceneDev& S, const RayPre& r, float tnear, float& tfar,
                                               uint32_t* stack)
{
    if constexpr (TRAV == kTravBinary)
        return trace<ANY>(S.nodes, S.tris, r, tnear, tfar, reinterpret_cast<int32_t*>(stack) + threadIdx.x, kTraceBlock);
    else if constexpr (TRAV == kTravWide)
        return trace4<ANY, kStack32>(S.nodes4, S.tris, r, tnear, tfar, stack, threadIdx.x, kTraceBlock);
    else if constexpr (TRAV == kTravWide16 || ANY)
        return trace4<ANY, kStack16>(S.nodes4, S.tris, r, tnear, tfar, stack, threadIdx.x, kTraceBlock);
    else if constexpr (TRAV == kTravWidePP)
        return trace4<ANY, kStackKeyed>(S.nodes4, S.tris, r, tnear, tfar, stack, threadIdx.x, kTraceBlock);
    else
        return trace4<ANY, kStackKeyed>(S.nodes4, S.tris, r, tnear, tfar, stack, threadIdx.x, kTraceBlock);
}

template <int TRAV>
constexpr int stack_kind() { return TRAV == kTravWide ? kStack32 : TRAV == kTravWide16 ? kStack16 : kStackKeyed; }

// Per-lane ray refill for the persistent 4-wide traversal kernels: a lane whose query is complete
// takes the next ray at once instead of waiting for the wavefront's slowest ray. The wavefront
// claims `chunk` queue entries at a time (its static share first, then through its partition's
// fetch cursor) and hands them to its idle lanes in lane order. Call with every lane of the
// wavefront; `busy` marks the lanes with a query.
// idle lanes that trigger a refill (unless the whole wavefront is idle): closest-hit and any-hit launches
// (with camera rays in their own packet kernel, 16 and 32 for either were 0.2-0.7 % slower than 24)
constexpr int kRefillMinExtend = 24;
constexpr int kRefillMinShadow = 24;
// Share of its even split of the queue that each wavefront takes without an atomic, in quarters
// (0: one fetch chunk). Larger static shares measured slower on the 82k-triangle scene, whose rays'
// costs vary across the image, and faster on the 22-triangle media scene, whose launches the
// fetch atomics dominated before the partitioned cursors.
constexpr uint32_t kStaticQuarters = 0;

struct RayRefill {
    uint32_t pos = 0, end = 0;  // the wavefront's claimed range of queue entries (uniform)
    uint32_t dyn_base = 0;      // first queue entry handed out through the fetch cursors
    uint32_t parts = 1, part = 0;
    bool drained = false;       // the queue has no entries left for this wavefront (uniform)

    // Every wavefront starts on its own contiguous static share of the queue (entries
    // [wave * share, (wave + 1) * share), at least one chunk) without an atomic. The entries after
    // all static shares are dealt in chunks, in queue order, through per-partition cursors: chunk
    // g = k * parts + p is the k-th fetch of partition p = blockIdx.x % parts, so the partitions'
    // wavefronts work through interleaved chunks (every partition sees the same mix of cheap and
    // costly rays) while their fetch atomics go to `parts` different lines. A launch with no
    // entries beyond the static shares takes no atomic at all (the last bounces of a render, whose
    // few rays otherwise waited on one failing fetch per wavefront of the grid). Handing chunks out
    // in queue order keeps the wavefronts on neighbouring rays (and so on the same part of the BVH
    // in cache).
    __device__ __forceinline__ RayRefill(uint32_t count, uint32_t chunk)
    {
        const uint32_t waves = gridDim.x * (blockDim.x >> 6);
        const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        uint32_t share = (count / waves) * kStaticQuarters / 4u;
        share = share < chunk ? chunk : share - share % chunk;
        const uint64_t first = uint64_t(wave) * share, all = uint64_t(waves) * share;
        pos = first < count ? uint32_t(first) : count;
        end = first + share < count ? uint32_t(first + share) : count;
        dyn_base = all < count ? uint32_t(all) : count;
        parts = gridDim.x < kParts ? gridDim.x : kParts;
        part = blockIdx.x % parts;
    }

    // Assigns queue entries to idle lanes; `start(i)` begins the query of entry i on this lane.
    // `cursor` is the partition word (P_FETCH_EXT / P_FETCH_SHADOW) of the launch's queue.
    template <int kRefillMin = kRefillMinExtend, class Start>
    __device__ __forceinline__ void run(uint32_t count, uint32_t chunk, uint32_t* counters, uint32_t cursor, bool& busy,
                                        Start start)
    {
        const uint32_t lane = __lane_id();
        const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
        uint64_t need = __ballot(!busy);
        // refill in bulk: ray set-up runs once for many idle lanes instead of once per finished ray
        // (24 idle lanes measured best on the 82k-triangle scene: 1-2 % faster than 16 or 32, 4 % than 48,
        // and 13 % faster than refilling each lane at once)
        if (__popcll(need) < kRefillMin && need != ~0ull) return;
        while (need && !drained) {
            if (pos >= end) {
                if (dyn_base >= count) {  // nothing beyond the static shares
                    drained = true;
                    break;
                }
                uint32_t k = 0;
                if (lane == 0) k = atomicAdd(counters + part_word(part, cursor), 1u);
                k = __shfl(k, 0);
                const uint64_t b = uint64_t(dyn_base) + (uint64_t(k) * parts + part) * chunk;
                if (b >= count) {
                    drained = true;
                    break;
                }
                pos = uint32_t(b);
                end = uint32_t(min(b + chunk, uint64_t(count)));
            }
            const uint32_t avail = end - pos;
            const uint32_t rank = uint32_t(__popcll(need & below));
            if (!busy && rank < avail) {
                start(pos + rank);
                busy = true;
            }
            pos += min(uint32_t(__popcll(need)), avail);
            need = __ballot(!busy);
        }
    }
};

// ---------------------------------------------------------------------------------------------
// extend: closest hit for each queued ray (7 wavefronts per SIMD: the batched variant's per-lane
// scene pointers fit the 72-VGPR budget without spills)
template <int TRAV, int STACK, bool BATCH>
__global__ void __launch_bounds__(kTraceBlock, 7) extend_kernel(SceneDev S0, const JobDev* __restrict__ jobs, WaveDev Wv,
                                                             int cur_first)
{
    __shared__ uint32_t stack[stack_words<TRAV, STACK>() * kTraceBlock];
    const int cur = cur_first & 1;
    const bool first = (cur_first & 2) != 0;  // camera rays: the origin is the camera's
    const uint32_t count = Wv.counters[qword(cur)];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        Wv.counters[qword(cur ^ 1)] = 0;
        Wv.counters[shword(cur ^ 1)] = 0;
        atomicAdd(reinterpret_cast<unsigned long long*>(Wv.counters + C_STAT_EXT), (unsigned long long)count);
    }
    if (gid < kParts) Wv.counters[part_word(gid, P_FETCH_SHADOW)] = 0;
    const uint32_t chunk = fetch_chunk(count);
    if constexpr (TRAV == kTravWidePP) {
        // persistent grid with per-lane refill; node steps and triangle steps in wavefront-wide phases
        RayRefill rf(count, chunk);
        bool busy = false, blocked = false;
        uint32_t rec = 0;
        RayPre r;
        Trav4 st;
        const float4 *nodes = S0.nodes4, *tris = S0.tris;
        for (;;) {
            rf.run(count, chunk, Wv.counters, P_FETCH_EXT, busy, [&](uint32_t i) {
                rec = i;
                const float4 d = Wv.ray_d[cur][i];
                float3 c = Wv.cpos;
                float pb = S0.plane_bound;
                if constexpr (BATCH) {
                    const JobDev& J = jobs[__float_as_uint(d.w)];
                    nodes = J.S.nodes4;
                    tris = J.S.tris;
                    c = J.R.cpos;
                    pb = J.S.plane_bound;
                }
                const float4 o = first ? make_float4(c.x, c.y, c.z, 0.0f) : Wv.ray_o[cur][i];
                r = ray_pre(mk(o), mk(d), pb);
                trav4_begin(st, __int_as_float(0x7f800000));
                blocked = false;
            });
            if (__ballot(busy) == 0ull) break;  // the queue is drained and every query complete
            const bool can = busy && trav4_pp_can_step(st, blocked);
            const bool pend = busy && st.leaves != 0u;
            const uint64_t cm = __ballot(can), pm = __ballot(pend);
            bool done = busy && !can && !pend;  // (every busy lane can step or holds leaves)
            if (cm == 0ull || __popcll(pm) >= kLeafPhaseMin) {
                if (pend) done = trace4_leaf_pp(nodes, tris, r, kSigmaVisibility, st, blocked);
            } else if (can) {
                done = trace4_node_pp<kStackKeyed>(nodes, r, kSigmaVisibility, st, stack, threadIdx.x, kTraceBlock,
                                                   blocked);
            }
            if (busy && done) {
                Wv.hit[rec] = make_float2(st.hit >= 0 ? st.tfar : -1.0f, __int_as_float(st.hit));
                busy = false;
            }
        }
    } else if constexpr (TRAV != kTravBinary) {
        // persistent grid with per-lane refill: every lane steps its own query one node at a time
        RayRefill rf(count, chunk);
        bool busy = false;
        uint32_t rec = 0;
        RayPre r;
        Trav4 st;
        const float4 *nodes = S0.nodes4, *tris = S0.tris;
        for (;;) {
            rf.run(count, chunk, Wv.counters, P_FETCH_EXT, busy, [&](uint32_t i) {
                rec = i;
                const float4 d = Wv.ray_d[cur][i];
                float3 c = Wv.cpos;
                float pb = S0.plane_bound;
                if constexpr (BATCH) {
                    const JobDev& J = jobs[__float_as_uint(d.w)];
                    nodes = J.S.nodes4;
                    tris = J.S.tris;
                    c = J.R.cpos;
                    pb = J.S.plane_bound;
                }
                const float4 o = first ? make_float4(c.x, c.y, c.z, 0.0f) : Wv.ray_o[cur][i];
                r = ray_pre(mk(o), mk(d), pb);
                trav4_begin(st, __int_as_float(0x7f800000));
            });
            if (__ballot(busy) == 0ull) break;  // the queue is drained and every query complete
            if (busy && trace4_step<false, stack_kind<TRAV>()>(nodes, tris, r, kSigmaVisibility, st, stack, threadIdx.x,
                                                                kTraceBlock)) {
                Wv.hit[rec] = make_float2(st.hit >= 0 ? st.tfar : -1.0f, __int_as_float(st.hit));
                busy = false;
            }
        }
    } else {
        // persistent grid: each wavefront fetches `chunk` queue entries at a time until the queue
        // is drained (one fetch atomic per `chunk` rays)
        const uint32_t lane = __lane_id();
        for (;;) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(Wv.counters + part_word(0, P_FETCH_EXT), chunk);
            base = __shfl(base, 0);
            if (base >= count) break;
            for (uint32_t j = 0; j < chunk; j += 64) {
                const uint32_t i = base + j + lane;
                if (i < count) {
                    const float4 d = Wv.ray_d[cur][i];
                    const SceneDev& S = BATCH ? jobs[__float_as_uint(d.w)].S : S0;
                    const float3 c = BATCH ? jobs[__float_as_uint(d.w)].R.cpos : Wv.cpos;
                    const float4 o = first ? make_float4(c.x, c.y, c.z, 0.0f) : Wv.ray_o[cur][i];
                    RayPre r = ray_pre(mk(o), mk(d), S.plane_bound);
                    float tfar = __int_as_float(0x7f800000);
                    int32_t tri = trace_scene<false, TRAV>(S, r, kSigmaVisibility, tfar, stack);
                    Wv.hit[i] = make_float2(tri >= 0 ? tfar : -1.0f, __int_as_float(tri));
                }
            }
        }
    }
}

// extend for the camera rays of single-job renders (the first bounce): the queue holds them in path
// order, so a wavefront's 64 consecutive entries are samples of one pixel, traced as one packet
// (trace4_packet); the lanes finish together and the refill hands them the next 64 entries. The
// hits are those of extend_kernel (the flag TR_RENDER_NO_CAMERA_PACKETS selects it instead).
template <int TRAV, int STACK>
__global__ void __launch_bounds__(kTraceBlock, 7) extend_camera_kernel(SceneDev S0, WaveDev Wv, int cur)
{
    static_assert(TRAV == kTravWide || TRAV == kTravWide16 || TRAV == kTravWideKeyed, "4-wide stacks");
    __shared__ uint32_t stack[stack_words<TRAV, STACK>() * kTraceBlock];
    const uint32_t count = Wv.counters[qword(cur)];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        Wv.counters[qword(cur ^ 1)] = 0;
        Wv.counters[shword(cur ^ 1)] = 0;
        atomicAdd(reinterpret_cast<unsigned long long*>(Wv.counters + C_STAT_EXT), (unsigned long long)count);
    }
    if (gid < kParts) Wv.counters[part_word(gid, P_FETCH_SHADOW)] = 0;
    const uint32_t chunk = fetch_chunk(count);
    RayRefill rf(count, chunk);
    bool busy = false;
    uint32_t rec = 0;
    RayPre r;
    const int col = int(threadIdx.x & ~63u);
    const f3 cam = mk(Wv.cpos.x, Wv.cpos.y, Wv.cpos.z);
    for (;;) {
        rf.run(count, chunk, Wv.counters, P_FETCH_EXT, busy, [&](uint32_t i) {
            rec = i;
            r = ray_pre(cam, mk(Wv.ray_d[cur][i]), S0.plane_bound);
        });
        if (__ballot(busy) == 0ull) break;  // the queue is drained
        float tfar = __int_as_float(0x7f800000);
        int32_t hit = -1;
        trace4_packet<stack_kind<TRAV>()>(S0.nodes4, S0.tris, r, kSigmaVisibility, busy, tfar, hit, stack, col,
                                          kTraceBlock);
        if (busy) {
            Wv.hit[rec] = make_float2(hit >= 0 ? tfar : -1.0f, __int_as_float(hit));
            busy = false;
        }
    }
}

// Lambertian cosine-weighted direction in the reference's shading frame
// (Sampling.h:131-143, Intersection.h:167-178, Lambertian.h:107-122).
__device__ __forceinline__ f3 sample_lambert(f3 n, float e1, float e2, float* cos_sampled)
{
    float phi = 2.0f * kPi * e1;
    float ct = sqrtf(e2);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    f3 l = mk(cp * st, ct, sp * st);
    f3 aux = (fabsf(n.x) > fabsf(n.z)) ? mk(-n.y, n.x, n.z) : mk(0.0f, -n.z, n.y);
    f3 tv = normalize(aux);
    f3 tu = cross(n, tv);
    *cos_sampled = ct;
    return normalize(tu * l.x + n * l.y + tv * l.z);
}

// Cosine-weighted direction in an orthonormal frame around n (Duff et al. 2017); the density is
// exactly cos/pi, which one-sample MIS with the guiding mixture needs.
__device__ __forceinline__ f3 sample_lambert_onb(f3 n, float e1, float e2, float* cos_sampled)
{
    float phi = 2.0f * kPi * e1;
    float ct = sqrtf(e2);
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    f3 t = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    f3 bt = mk(b, sign + n.y * n.y * a, -n.y);
    *cos_sampled = ct;
    return normalize(t * (cp * st) + bt * (sp * st) + n * ct);
}

// Interpolated shading normal of a smooth triangle at point p (per-corner normals n0, n1, n2 with
// the barycentrics of p), not renormalised, as the reference uses it (World.h:589-598,
// Triangle.h:155-159).
__device__ __forceinline__ f3 smooth_normal(const float4* V, f3 p)
{
    f3 p0 = mk(V[0]), e1 = mk(V[1]), e2 = mk(V[2]), q = p - p0;
    float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2), d20 = dot(q, e1), d21 = dot(q, e2);
    float inv = 1.0f / (d00 * d11 - d01 * d01);
    float u = (d11 * d20 - d01 * d21) * inv, v = (d00 * d21 - d01 * d20) * inv;
    return mk(V[3]) * (1.0f - u - v) + mk(V[4]) * u + mk(V[5]) * v;
}

// tVectorN::transform_matrix_to(from = (0,1,0), to) applied to v (VectorN.h:429-523): the
// rotation taking +y onto `to`, built from `to` as given (the time-angular sampler passes an
// unnormalised vector, as the reference does), result normalised.
__device__ __forceinline__ f3 rotate_from_y(f3 v, f3 wi)
{
    f3 from = mk(0.0f, 1.0f, 0.0f);
    f3 vx = cross(from, wi);
    float e = wi.y;
    float f = fabsf(e);
    float m[3][3];
    if (double(f) > 1.0 - double(1.e-4f)) {
        f3 x = mk(1.0f, 0.0f, 0.0f);  // smallest |component| of (0,1,0) with the reference's tie order
        f3 u = x - from, w = x - wi;
        float c1 = 2.0f / dot(u, u), c2 = 2.0f / dot(w, w), c3 = c1 * c2 * dot(u, w);
        float ua[3] = {u.x, u.y, u.z}, wa[3] = {w.x, w.y, w.z};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = 0; j < 3; ++j) m[i][j] = -c1 * ua[i] * ua[j] - c2 * wa[i] * wa[j] + c3 * wa[i] * ua[j];
            m[i][i] += 1.0f;
        }
    } else {
        float h = 1.0f / (1.0f + e);
        float hvx = h * vx.x, hvz = h * vx.z;
        float hvxy = hvx * vx.y, hvxz = hvx * vx.z, hvyz = hvz * vx.y;
        m[0][0] = e + hvx * vx.x;
        m[0][1] = hvxy - vx.z;
        m[0][2] = hvxz + vx.y;
        m[1][0] = hvxy + vx.z;
        m[1][1] = e + h * vx.y * vx.y;
        m[1][2] = hvyz - vx.x;
        m[2][0] = hvxz - vx.y;
        m[2][1] = hvyz + vx.x;
        m[2][2] = e + hvz * vx.z;
    }
    f3 s = mk(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
    return normalize(s);
}

// Isotropic phase-function direction: uniform sphere rotated so that +y maps onto wi
// (Sampling.h:80-94, IsotropicPhaseFunction.h:54-60); the new Ray normalises it again.
__device__ __forceinline__ f3 sample_isotropic(f3 wi, float e1, float e2)
{
    float phi = 2.0f * kPi * e1;
    float ct = 1.0f - 2.0f * e2;
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return normalize(rotate_from_y(mk(cp * st, ct, sp * st), wi));
}

// Time-angular direction sampling at a medium vertex (AngularSampling::ShortestTime,
// BidirectionalPathTracing.h:357-398 with Probability::TimeAngular, SamplingTechniques.h:170-250;
// eta = 1, c = 1/time_of_flight(1) = 1): the angle theta to the direction of the light point `l`
// is drawn so that a free flight of r followed by the straight line to the light arrives
// uniformly in the time window [ta, tb] (clipped to what the geometry allows), ta being the
// straight-line arrival time. The caller first checks time_angular_window (no draw): when the window
// left after this flight is shorter than one exposure it samples the phase function. Returns false
// when the density falls below 1e-5 (the caller then samples the phase function with two more
// draws). pdf is per (theta, phi), as the reference uses it.
__device__ __forceinline__ bool time_angular_window(float modl, float time, float r, float tb, float exposure)
{
#pragma clang fp contract(off)
    const float ta = modl * 1.0f + time;
    return !((tb - (ta + r * 1.0f)) < exposure);
}

__device__ __forceinline__ bool sample_time_angular(f3 l, float modl, float time, float r, float tb, float e_theta,
                                                    float e_phi, f3* dir, float* pdf)
{
#pragma clang fp contract(off)
    const float ta = modl * 1.0f + time;
    // set_temporal_boundaries with ec = 1
    float t_a = r > modl ? fmaxf(ta, 2.0f * r - modl) : fmaxf(ta, modl);
    float t_b = fminf(tb, 2.0f * r + modl);
    float t = e_theta * (t_b - t_a) + t_a;
    float cos_t = float((double(modl * modl) + 2.0 * double(t) * double(r) - double(t * t)) /
                        (2.0 * double(r) * double(modl)));
    cos_t = fmaxf(-1.0f, fminf(cos_t, 1.0f));
    const float theta = float(acos(double(cos_t)));  // correctly rounded, as the host's acosf
    const float phi = float(2.0 * 3.14159265358979323846 * double(e_phi));
    float sp, cp, st, ct;
    sincosf(phi, &sp, &cp);
    sincosf(theta, &st, &ct);
    *dir = normalize(rotate_from_y(mk(cp * st, ct, sp * st), l));  // the new Ray normalises
    const double den = sqrt(double(r * r + modl * modl) - 2.0 * double(r) * double(modl) * double(ct));
    const float dtdtheta = float(double(modl * r * st) / den);
    *pdf = float(double(dtdtheta / (t_b - t_a)) / (2.0 * 3.14159265358979323846));
    return !(double(*pdf) < 1e-5);
}

// The light point a path aims at in the Transient sampling mode: the reference draws it per
// incoming sample with World::sample_light + LightSource::sample(ls, pdf)
// (BidirectionalPathTracing.h:1115-1118; PointLightSource.h:173-183, RectangularAreaLightSource.h:
// 250-271); here it comes from the path's own stream seq = 3 * 2^61 | path id.
__device__ __forceinline__ f3 light_target(const SceneDev& S, uint64_t seed, uint64_t pid)
{
#pragma clang fp contract(off)
    const uint64_t seq = (3ull << 61) | pid;
    uint64_t tinc = pcg_inc(seq), tst = pcg_seed(seed, seq);
    float u = pcg_float(tst, tinc);
    int li = min(int(u * float(S.n_lights)), S.n_lights - 1);
    const LightHost& L = S.lights[li];
    f3 p = mk(L.pos[0], L.pos[1], L.pos[2]);
    if (int(L.I[3]) != kLightPoint) {
        float a = pcg_float(tst, tinc), b = pcg_float(tst, tinc);
        p = mk((p.x + L.left[0] * a) + L.up[0] * b, (p.y + L.left[1] * a) + L.up[1] * b,
               (p.z + L.left[2] * a) + L.up[2] * b);
    }
    return p;
}

// ---------------------------------------------------------------------------------------------
// shade: one thread per queued path whose ray has just been traced
// Shading modes: the reference's shading frame (default), an orthonormal frame (same estimator
// without the frame quirk), and guided sampling (orthonormal frame + one-sample MIS of the BSDF /
// phase function with the network's vMF mixture).
enum { kShadeRef = 0, kShadeOnb = 1, kShadeGuided = 2 };
constexpr int kShadeBlockGuided = 256;  // 4 wavefronts x 10 KiB of LDS for the network

__device__ __forceinline__ float mean3(f3 v) { return (v.x + v.y + v.z) * (1.0f / 3.0f); }

template <int MODE, bool BATCH>
__global__ void __launch_bounds__(MODE == kShadeGuided ? kShadeBlockGuided : kShadeBlock)
    shade_kernel(SceneDev S0, RenderDev R0, const JobDev* __restrict__ jobs, WaveDev Wv, int cur, trg::GuideView G)
{
    static_assert(!(BATCH && MODE == kShadeGuided), "guided renders take one job");
    constexpr bool GUIDED = MODE == kShadeGuided;
    __shared__ uint32_t scan[kShadeBlock / 64 + 2];
    __shared__ float glds[GUIDED ? (kShadeBlockGuided / 64) * 64 * trg::kRow : 1];
    __shared__ uint4 gfrag[GUIDED ? trg::kFwdFrags * 64 : 1];
    const uint32_t count = Wv.counters[qword(cur)];
    if (blockIdx.x == 0 && threadIdx.x < kParts) Wv.counters[part_word(threadIdx.x, P_FETCH_EXT)] = 0;  // for the next extend
    if (blockIdx.x * blockDim.x >= count) return;  // no paths for this block (the last bounces' launches)
    if constexpr (GUIDED) {
        // the network's weight fragments stay in LDS for the whole block (38 KiB; with the rows of
        // four wavefronts 80.3 KB, two blocks per CU); the biases are read from global memory
        for (uint32_t k = threadIdx.x; k < uint32_t(trg::kFwdFrags * 64); k += blockDim.x) gfrag[k] = G.frag[k];
        __syncthreads();
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride) {
        const uint32_t i = base + threadIdx.x;
        bool live = i < count;
        const float4 d4 = live ? Wv.ray_d[cur][i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // batched renders: the path's job; otherwise its index in the wavefront
        const uint32_t slot = __float_as_uint(d4.w);
        const uint32_t jid = BATCH ? slot : 0u;
        const SceneDev& S = BATCH ? jobs[jid].S : S0;
        const RenderDev& R = BATCH ? jobs[jid].R : R0;
        bool emit_shadow = false, cont = false, want = false;
        float4 sh_o = {}, sh_d = {}, sh_c = {};
        float4 n_o = {}, n_d = {}, n_thr = {};
        uint4 n_misc = {}, n_extra = {};
        uint32_t sh_tag = 0xFFFFFFFFu, tag = 0xFFFFFFFFu;
        // state of a vertex that goes on to sample a direction
        f3 f = {}, pos = {}, n = {}, d = {};
        bool medium_event = false, two_sided = false;
        int32_t mat = -1;
        float pi = 1.0f, time = 0.0f, nsampled = 0.0f;
        uint64_t pst = 0, nst = 0, pinc = 0;
        uint4 misc = {}, extra = {};
        if (live) {
            misc = Wv.misc[cur][i];
            extra = Wv.extra[cur][i];
            float2 h = Wv.hit[i];
            int32_t tri = __float_as_int(h.y);
            uint32_t level = extra.z & 0xFFu;
            // a camera ray starts at the camera with unit throughput; extra.w holds its sampled
            // free-flight distance until the unwarp distance replaces it below
            const float4 o4 = level == 0 ? make_float4(R.cpos.x, R.cpos.y, R.cpos.z, __uint_as_float(extra.w))
                                         : Wv.ray_o[cur][i];
            const float4 thr = level == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : Wv.thr[cur][i];
            f3 o = mk(o4);
            d = mk(d4);
            float sampled = o4.w;
            if (level == 0 && R.unwarp) extra.w = __float_as_uint(tri >= 0 ? h.x : 3.402823466e+38f);
            if constexpr (GUIDED) {
                if (slot % G.stride == 0u && slot / G.stride < G.train_paths) tag = (slot / G.stride) * G.max_levels;
            }
            // correct_next_event_distance (BidirectionalPathTracing.h:275-307)
            float dist, pd = 1.0f;
            bool alive = true;
            const float lam = R.time_sampling ? R.ts_lambda : S.lambda;  // NextEventSampling::Time / MeanFreePath
            if (!S.has_medium) {
                alive = tri >= 0;
                dist = h.x;
            } else if (tri >= 0 && h.x < sampled) {
                dist = h.x;
                pd = expf(-dist * lam);
            } else {
                dist = sampled;
                medium_event = true;
                pd = expf(-dist * lam) * lam;
            }
            if (alive) {
                f = mk(thr);
                if (S.has_medium) {
                    f = mk(f.x * expf(-dist * S.sigma_t.x), f.y * expf(-dist * S.sigma_t.y), f.z * expf(-dist * S.sigma_t.z));
                    f = f * (1.0f / pd);
                }
                time = thr.w + dist;
                pos = point_on_ray(o, d, dist);
                n = mk(0.0f, 0.0f, 0.0f);
                if (!medium_event) {
                    float4 nn = S.tri_normal[tri];
                    n = mk(nn);
                    mat = __float_as_int(nn.w);
                    two_sided = (mat & kMatTwoSided) != 0;
                    if (mat & kMatSmooth) n = smooth_normal(S.tri_vn + 6 * size_t(tri), pos);
                    mat &= kMatIndex;
                }
                nst = u64(extra.x, extra.y);
                uint64_t pid = u64(misc.x, misc.y);
                uint64_t ninc = pcg_inc(2 * pid + 1);
                pinc = pcg_inc(2 * pid);
                // ---- next-event estimation at this vertex (connect_paths :976-1045,
                //      connect_vertex_with_light :650-692)
                if (int(level) + 1 <= R.max_bounces && time <= R.total_time) {
                    float u = pcg_float(nst, ninc);
                    int li = min(int(u * float(S.n_lights)), S.n_lights - 1);
                    const LightHost& L = S.lights[li];
                    int type = int(L.I[3]);
                    f3 lp, ldir;
                    float ldist, pl2 = 1.0f;
                    f3 irr;
                    bool ok = true;
                    if (type == kLightPoint) {
                        lp = mk(L.pos[0], L.pos[1], L.pos[2]);
                        f3 v = pos - lp;
                        float d2 = dot(v, v);
                        ldist = sqrtf(d2);
                        ldir = v * (1.0f / ldist);
                        float att = 1.0f / d2;
                        irr = mk(L.I[0] * att, L.I[1] * att, L.I[2] * att);
                    } else {
                        float uu = pcg_float(nst, ninc), vv = pcg_float(nst, ninc);
                        lp = mk(L.pos[0], L.pos[1], L.pos[2]) + mk(L.left[0], L.left[1], L.left[2]) * uu +
                             mk(L.up[0], L.up[1], L.up[2]) * vv;
                        f3 v = pos - lp;
                        ldist = sqrtf(dot(v, v));
                        ldir = v * (1.0f / ldist);
                        float dl = dot(ldir, mk(L.normal[0], L.normal[1], L.normal[2]));
                        ok = dl > 0.0f;
                        float a = L.I[0] * dl / (ldist * ldist);
                        irr = mk(a, a, a);
                        pl2 = L.normal[3];
                        ldir = normalize(ldir);
                    }
                    if (ok) {
                        f3 fe;
                        if (medium_event) {
                            fe = mk(S.sigma_s.x, S.sigma_s.y, S.sigma_s.z) * kInv4Pi;
                        } else {
                            float4 kd = S.mats[mat];
                            // dot_clamped, or dot_abs for a two-sided material (Path.h:188-193)
                            float cd = -dot(ldir, n);
                            float c = two_sided ? fabsf(cd) : fmaxf(cd, 0.0f);
                            fe = mk(kd.x, kd.y, kd.z) * (kInvPi * c);
                        }
                        f3 att = mk(1.0f, 1.0f, 1.0f);
                        if (S.has_medium)
                            att = mk(expf(-S.sigma_t.x * ldist), expf(-S.sigma_t.y * ldist), expf(-S.sigma_t.z * ldist));
                        float scale = R.inv_inc / (pl2 / float(S.n_lights));
                        f3 val = mul(mul(f, fe), mul(irr, att)) * scale;
                        float ts = ldist + time;
                        float cam_time = R.unwarp ? ts - __uint_as_float(extra.w) : ts;
                        float sensor = (cam_time - R.offset) / R.exposure;
                        bool finite = isfinite(val.x) && isfinite(val.y) && isfinite(val.z);
                        bool nonzero = val.x != 0.0f || val.y != 0.0f || val.z != 0.0f;
                        const bool in_window = sensor >= 0.0f && sensor < float(R.T);
                        if (!BATCH && !GUIDED && R.layers > 0) {
                            // multibounce film: every connected sample counts for its layer's
                            // weight and the steady image, in the time window or not
                            // (MultibounceStreakCameraFilm.h:137-157, Film.h:263-283)
                            const uint32_t layer = min(level, uint32_t(R.layers - 1));
                            uint32_t fidx = 0xFFFFFFFFu;
                            if (in_window)
                                fidx = layer * R.layer_stride + (extra.z >> 8) * uint32_t(R.T) +
                                       min(uint32_t(floorf(sensor)), uint32_t(R.T - 1));
                            if (!finite) val = mk(0.0f, 0.0f, 0.0f);  // TransientRenderer.h:167-171
                            emit_shadow = true;
                            sh_o = make_float4(lp.x, lp.y, lp.z, ldist);
                            sh_d = make_float4(ldir.x, ldir.y, ldir.z, __uint_as_float(fidx));
                            sh_c = make_float4(val.x, val.y, val.z, __uint_as_float(extra.z >> 8));
                            sh_tag = ((slot / uint32_t(R.n_inc)) << kMaxLayersTagBits) | layer;
                        } else if (in_window && finite && nonzero) {
                            uint32_t bin = min(uint32_t(floorf(sensor)), uint32_t(R.T - 1));
                            uint32_t fidx = (extra.z >> 8) * uint32_t(R.T) + bin;
                            emit_shadow = true;
                            sh_o = make_float4(lp.x, lp.y, lp.z, ldist);
                            sh_d = make_float4(ldir.x, ldir.y, ldir.z, __uint_as_float(fidx));
                            sh_c = make_float4(val.x, val.y, val.z, __uint_as_float(extra.z >> 8));
                            if (GUIDED && tag != 0xFFFFFFFFu) sh_tag = tag + level;
                        }
                    }
                }
                // ---- termination: later vertices cannot contribute past the bounce cap or the
                //      film window; otherwise Russian roulette on the mean albedo, or in the
                //      Transient mode the film's time length alone (TotalTime) (:447-481)
                if (int(level) + 1 < R.max_bounces && time <= R.total_time) {
                    pst = u64(misc.z, misc.w);
                    float albedo = medium_event ? S.medium_albedo : S.mats[mat].w;
                    float eps = R.time_sampling ? 0.0f : pcg_float(pst, pinc);
                    if (!(eps > albedo)) {
                        pi = R.time_sampling ? 1.0f : albedo;
                        nsampled = __int_as_float(0x7f800000);
                        if (S.has_medium) nsampled = free_flight(pcg_float(pst, pinc), lam);
                        want = true;
                        if (GUIDED && tag != 0xFFFFFFFFu) tag += level;
                    }
                }
            }
        }
        if (!want) tag = 0xFFFFFFFFu;
        // ---- guided: the network's mixture for every vertex of this wavefront that samples a
        //      direction (all 64 lanes take part in the matrix-core evaluation)
        float raw[GUIDED ? trg::kRaw : 1];
        float* mine = glds + (GUIDED ? (threadIdx.x >> 6) * 64 * trg::kRow : 0);
        if constexpr (GUIDED) {
            const f3 pn = mk((pos.x - G.bb_lo.x) * G.bb_inv.x, (pos.y - G.bb_lo.y) * G.bb_inv.y,
                             (pos.z - G.bb_lo.z) * G.bb_inv.z);
            raw[0] = pn.x, raw[1] = pn.y, raw[2] = pn.z;
            raw[3] = -d.x, raw[4] = -d.y, raw[5] = -d.z;
            raw[6] = n.x, raw[7] = n.y, raw[8] = n.z;
            raw[9] = fminf(fmaxf(time / R.total_time, 0.0f), 1.0f);
            raw[10] = medium_event ? 1.0f : 0.0f;
            raw[11] = 0.0f;
            if (__ballot(want) != 0ull) {
                const int lane = int(__lane_id());
                float feat[trg::kFeat];
                trg::encode_features(raw, feat);
                trg::store_feature_row(reinterpret_cast<uint32_t*>(mine) + lane * trg::kRow, feat);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                trg::mlp_forward_wave<1>(mine, gfrag, G.bias, nullptr);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        if (want) {
            f3 nd, fa;
            float pa;
            bool ok = true;
            if constexpr (!GUIDED) {
                if (medium_event) {
                    bool timed = false;
                    if (R.time_sampling) {
                        const f3 tg = light_target(S, R.seed, u64(misc.x, misc.y));
                        f3 l;
                        float modl;
                        {
#pragma clang fp contract(off)
                            l = mk(tg.x - pos.x, tg.y - pos.y, tg.z - pos.z);
                            modl = sqrtf((l.x * l.x + l.y * l.y) + l.z * l.z);
                        }
                        if (time_angular_window(modl, time, nsampled, R.total_time, R.exposure)) {
                            float e1 = pcg_float(pst, pinc), e2 = pcg_float(pst, pinc);
                            timed = sample_time_angular(l, modl, time, nsampled, R.total_time, e1, e2, &nd, &pa);
                        }
                    }
                    if (!timed) {
                        float e1 = pcg_float(pst, pinc), e2 = pcg_float(pst, pinc);
                        nd = sample_isotropic(d, e1, e2);
                        pa = kInv4Pi;
                    }
                    fa = mk(S.sigma_s.x, S.sigma_s.y, S.sigma_s.z) * kInv4Pi;
                } else {
                    (void)pcg_float(pst, pinc);  // BSDF pick of a single-BSDF UberMaterial
                    float e1 = pcg_float(pst, pinc), e2 = pcg_float(pst, pinc);
                    float ct;
                    nd = MODE == kShadeRef ? sample_lambert(n, e1, e2, &ct) : sample_lambert_onb(n, e1, e2, &ct);
                    // a two-sided material hit from behind samples in its inverted frame
                    // (TwoSided.h:80-88, Intersection.h:180-185)
                    if (two_sided && dot(d, n) > 0.0f) nd = nd * -1.0f;
                    float4 kd = S.mats[mat];
                    fa = mk(kd.x, kd.y, kd.z) * (kInvPi * fabsf(dot(n, nd)));
                    pa = ct * kInvPi;
                }
            } else {
                const float alpha = G.alpha;
                const f3 nside = (two_sided && dot(d, n) > 0.0f) ? n * -1.0f : n;  // side the path arrives on
                float usel = pcg_float(pst, pinc);
                float e1, e2;
                if (usel < alpha) {
                    e1 = pcg_float(pst, pinc), e2 = pcg_float(pst, pinc);
                    float ct;
                    nd = medium_event ? sample_isotropic(d, e1, e2) : sample_lambert_onb(nside, e1, e2, &ct);
                }
                trg::Mixture mix;
                trg::mixture_decode(mine + __lane_id() * trg::kRow, mix);
                if (!(usel < alpha)) {
                    float ul = pcg_float(pst, pinc);
                    e1 = pcg_float(pst, pinc), e2 = pcg_float(pst, pinc);
                    trg::Lobe L = trg::lobe_of(mix, trg::pick_lobe(mix, ul));
                    trg::vmf_sample(L, e1, e2, &nd.x, &nd.y, &nd.z);
                }
                float q = trg::mixture_pdf(mix, nd.x, nd.y, nd.z);
                if (medium_event) {
                    fa = mk(S.sigma_s.x, S.sigma_s.y, S.sigma_s.z) * kInv4Pi;
                    pa = alpha * kInv4Pi + (1.0f - alpha) * q;
                } else {
                    float c = dot(nside, nd);
                    ok = c > 0.0f;  // below the surface the one-sided Lambertian reflects nothing
                    float4 kd = S.mats[mat];
                    fa = mk(kd.x, kd.y, kd.z) * (kInvPi * fmaxf(c, 0.0f));
                    pa = alpha * fmaxf(c, 0.0f) * kInvPi + (1.0f - alpha) * q;
                }
                ok = ok && pa > 0.0f;
                if (ok && tag != 0xFFFFFFFFu) {
                    // training record: inputs, sampled direction and pi / throughput (the later
                    // contributions along this direction, divided by the throughput after this
                    // vertex, times BSDF x cosine / density)
                    float mf = mean3(f);
                    float4* rec = G.rec + size_t(tag) * 4;
                    rec[0] = make_float4(raw[0], raw[1], raw[2], raw[9]);
                    rec[1] = make_float4(raw[3], raw[4], raw[5], raw[10]);
                    rec[2] = make_float4(raw[6], raw[7], raw[8], mf > 0.0f ? pi / mf : 0.0f);
                    rec[3] = make_float4(nd.x, nd.y, nd.z, 0.0f);
                    G.n_rec[tag / G.max_levels] = tag % G.max_levels + 1u;
                }
            }
            if (ok) {
                f3 nf = mul(f, fa) * (1.0f / (pi * pa));
                cont = true;
                n_o = make_float4(pos.x, pos.y, pos.z, nsampled);
                n_d = make_float4(nd.x, nd.y, nd.z, d4.w);
                n_thr = make_float4(nf.x, nf.y, nf.z, time);
                n_misc = make_uint4(misc.x, misc.y, uint32_t(pst), uint32_t(pst >> 32));
                n_extra = make_uint4(uint32_t(nst), uint32_t(nst >> 32), extra.z + 1u, extra.w);
            }
        }
        uint32_t q_pos, sh_pos;
        block_append2(Wv.counters + qword(cur ^ 1), cont ? 1u : 0u, emit_shadow ? 1u : 0u, scan, &q_pos, &sh_pos);
        if (emit_shadow) {
            Wv.sh_o[sh_pos] = sh_o;
            Wv.sh_d[sh_pos] = sh_d;
            Wv.sh_c[sh_pos] = sh_c;
            if (GUIDED || (!BATCH && R.layers > 0)) Wv.sh_tag[sh_pos] = sh_tag;
            if constexpr (BATCH) Wv.sh_job[sh_pos] = jid;
        }
        if (cont) {
            Wv.ray_o[cur ^ 1][q_pos] = n_o;
            Wv.ray_d[cur ^ 1][q_pos] = n_d;
            Wv.thr[cur ^ 1][q_pos] = n_thr;
            Wv.misc[cur ^ 1][q_pos] = n_misc;
            Wv.extra[cur ^ 1][q_pos] = n_extra;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// shadow: occlusion test light -> vertex in (eps, dist - eps]; visible contributions go to the film
// (7 wavefronts per SIMD: the any-hit leaf mode's state fits 72 VGPRs without spills; at 64 VGPRs
// it spilled and ran 2 % slower)
template <int TRAV, int STACK, bool BATCH>
__global__ void __launch_bounds__(kTraceBlock, 7) shadow_kernel(SceneDev S0, const JobDev* __restrict__ jobs, WaveDev Wv,
                                                                             int sh)
{
    __shared__ uint32_t stack[stack_words<TRAV, STACK>() * kTraceBlock];
    const uint32_t count = Wv.counters[shword(sh)];
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(Wv.counters + C_STAT_SHADOW), (unsigned long long)count);
    const uint32_t lane = __lane_id();
    uint32_t visible = 0;
    const uint32_t chunk = fetch_chunk(count);
    if constexpr (TRAV != kTravBinary) {
        // persistent grid with per-lane refill (see extend_kernel)
        RayRefill rf(count, chunk);
        bool busy = false;
        uint32_t rec = 0;
        RayPre r;
        Trav4 st;
        const float4 *nodes = S0.nodes4, *tris = S0.tris;
        for (;;) {
            rf.run<kRefillMinShadow>(count, chunk, Wv.counters, P_FETCH_SHADOW, busy, [&](uint32_t i) {
                rec = i;
                float pb = S0.plane_bound;
                if constexpr (BATCH) {
                    const SceneDev& S = jobs[Wv.sh_job[i]].S;
                    nodes = S.nodes4;
                    tris = S.tris;
                    pb = S.plane_bound;
                }
                const float4 o = Wv.sh_o[i], d = Wv.sh_d[i];
                r = ray_pre(mk(o), mk(d), pb);
                trav4_begin(st, o.w - kSigmaVisibility);
            });
            if (__ballot(busy) == 0ull) break;
            // any hit: a keyed launch uses the 16-bit format (its LDS holds either)
            constexpr int SK = stack_kind<TRAV>() == kStackKeyed ? kStack16 : stack_kind<TRAV>();
            if (busy && trace4_step<true, SK>(nodes, tris, r, kSigmaVisibility, st, stack, threadIdx.x, kTraceBlock)) {
                const bool vis = st.hit < 0;
                Wv.sh_vis[rec] = vis ? 1 : 0;
                visible += vis ? 1u : 0u;
                busy = false;
            }
        }
    } else {
        for (;;) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(Wv.counters + part_word(0, P_FETCH_SHADOW), chunk);
            base = __shfl(base, 0);
            if (base >= count) break;
            for (uint32_t j = 0; j < chunk; j += 64) {
                const uint32_t i = base + j + lane;
                if (i < count) {
                    const SceneDev& S = BATCH ? jobs[Wv.sh_job[i]].S : S0;
                    float4 o = Wv.sh_o[i], d = Wv.sh_d[i];
                    RayPre r = ray_pre(mk(o), mk(d), S.plane_bound);
                    float tfar = o.w - kSigmaVisibility;
                    bool vis = trace_scene<true, TRAV>(S, r, kSigmaVisibility, tfar, stack) < 0;
                    Wv.sh_vis[i] = vis ? 1 : 0;
                    visible += vis ? 1u : 0u;
                }
            }
        }
    }
    visible = wave_sum(visible);
    if (lane == 0 && visible)
        atomicAdd(reinterpret_cast<unsigned long long*>(Wv.counters + part_word(blockIdx.x % kParts, P_STAT_CONTRIB)),
                  (unsigned long long)visible);
}

// ---------------------------------------------------------------------------------------------
// film accumulation with compensated sums. Every film element is a pair of FP32 accumulators
// (sum, comp). An addition of x is an atomic add that returns the old value a; the rounding error
// of the stored a + x is then recovered exactly with Knuth's TwoSum and added to comp, so after
// any number n of additions in any order sum + comp equals the exact total up to the rounding of
// the additions into comp: within (2u + O(n u^2)) sum|x| (u = 2^-24), against n u sum|x| for a plain
// FP32 running sum. The hardware atomic rounds to nearest-even like the vector ALU, which is what
// makes the recovered error exact.
__device__ __forceinline__ float two_sum_err(float a, float b, float s)
{
    // s = fl(a + b); returns (a + b) - s exactly
    const float bb = s - a;
    return (a - (s - bb)) + (b - bb);
}

// an RGB contribution into elements i, i + 1, i + 2 of the compensated pair (sum, comp) held in
// global memory: the three returning atomics are issued before the first result is awaited (one
// round trip instead of three)
__device__ __forceinline__ void film_add3(float* __restrict__ sum, float* __restrict__ comp, size_t i, float x,
                                          float y, float z)
{
    const float a0 = unsafeAtomicAdd(sum + i, x);
    const float a1 = unsafeAtomicAdd(sum + i + 1, y);
    const float a2 = unsafeAtomicAdd(sum + i + 2, z);
    const float e0 = two_sum_err(a0, x, a0 + x), e1 = two_sum_err(a1, y, a1 + y), e2 = two_sum_err(a2, z, a2 + z);
    if (e0 != 0.0f) unsafeAtomicAdd(comp + i, e0);
    if (e1 != 0.0f) unsafeAtomicAdd(comp + i + 1, e1);
    if (e2 != 0.0f) unsafeAtomicAdd(comp + i + 2, e2);
}

// a compensated partial (s, c) into the pair (sum[i], comp[i])
__device__ __forceinline__ void film_add_partial(float* __restrict__ sum, float* __restrict__ comp, size_t i, float s,
                                                 float c)
{
    if (s != 0.0f) {
        const float a = unsafeAtomicAdd(sum + i, s);
        c += two_sum_err(a, s, a + s);
    }
    if (c != 0.0f) unsafeAtomicAdd(comp + i, c);
}

// film: bins the visible contributions of this bounce into the film. A block takes a range of
// kFilmRange consecutive shadow records; their pixels are few (paths are enumerated pixel-major
// and queue compaction keeps runs of neighbouring paths together), so each block accumulates into
// LDS histograms of up to `pix_slots` pixels x T bins x RGB and flushes each non-zero bin with one
// compensated add. Records of pixels that find no slot try again after the flush, in a second
// round with the slots free, and only those left after kFilmRounds rounds go straight to HBM: the
// range's records come from a few append chunks of the shading kernel, each holding paths of one
// or two pixels, and the 256-path chunks of guided shading put more pixels in a range than the
// slots hold (the second round: film time -4 % on the 1024-spp icosphere frames, -7 % on guided
// media passes).
// The LDS histograms hold FP64 partials (non-returning LDS atomic adds): a partial of at most
// kFilmRange float contributions is exact to within 2^-44 of their magnitude sum, far inside the
// film's compensated FP32 bound, and is flushed as the float pair (hi, lo) = (float(p),
// float(p - hi)). Short ranges (512 records) measured best: 1.6x faster than 4096-record ranges on
// the 1024-bin films, whose per-pixel records spread over many bins. The kernel waits on its loads
// and returning atomics: 1024-thread blocks (two per CU, 32 wavefronts) ran it 12-16 % faster than
// 256-thread blocks (three per CU, 12 wavefronts), and 72 KB of LDS per block (three 1024-bin
// pixels instead of two, still two blocks per CU) 10-14 % faster again.
// (StreakCameraFilm::add_samples / add_temporal_sample, StreakCameraFilm.h:139-192, 280-297)
constexpr int kFilmBlock = 1024;
constexpr uint32_t kFilmRange = 512;
constexpr uint32_t kFilmLdsBytes = 72 * 1024;  // FP64 histograms: pix_slots * T * 3 * 8 bytes (3 pixels of 1024 bins)
constexpr uint32_t kFilmMaxSlots = 8;
constexpr int kFilmRounds = 2;
static_assert(kFilmRange <= uint32_t(kFilmBlock), "one record of a range per thread");

template <bool GUIDED>
__global__ void __launch_bounds__(kFilmBlock) film_kernel(RenderDev R, WaveDev Wv, int sh, uint32_t pix_slots, float* contrib)
{
    extern __shared__ double hist[];         // pix_slots * T * 3 partials
    __shared__ uint32_t slot_pix[kFilmMaxSlots];
    const uint32_t count = Wv.counters[shword(sh)];
    // guided renders: ranges of 256 records, about one append chunk of the guided shading kernel's
    // 256-thread blocks, so a range spans fewer pixels (guided passes' film time -11 %)
    constexpr uint32_t kRange = GUIDED ? 256u : kFilmRange;
    if (blockIdx.x * kRange >= count) return;  // no range for this block (the last bounces' launches)
    const uint32_t T = uint32_t(R.T), T3 = 3u * T, hsize = pix_slots * T3;
    const int rounds = pix_slots ? kFilmRounds : 1;
    // the histograms start at zero; each flush reads and clears only the slots that were used
    for (uint32_t i = threadIdx.x; i < hsize; i += blockDim.x) hist[i] = 0.0;
    for (uint32_t start = blockIdx.x * kRange; start < count; start += gridDim.x * kRange) {
        // this thread's record of the range, if it is a visible contribution
        const uint32_t r = start + threadIdx.x;
        bool pending = false;
        float4 c = {};
        uint32_t pix = 0, bin = 0;
        if (threadIdx.x < kRange && r < count && Wv.sh_vis[r]) {
            c = Wv.sh_c[r];
            if constexpr (GUIDED) {
                // luminance of the contribution, credited to the training record of its vertex
                uint32_t tag = Wv.sh_tag[r];
                if (tag != 0xFFFFFFFFu) atomicAdd(contrib + tag, (c.x + c.y + c.z) * (1.0f / 3.0f));
            }
            pix = __float_as_uint(c.w);
            bin = __float_as_uint(Wv.sh_d[r].w) - pix * T;
            pending = true;
        }
        for (int round = 0; round < rounds; ++round) {
            if (threadIdx.x < kFilmMaxSlots) slot_pix[threadIdx.x] = 0xFFFFFFFFu;
            if (!__syncthreads_or(pending)) break;
            if (pending) {
                int slot = -1;
                for (uint32_t k = 0; k < pix_slots; ++k) {
                    uint32_t v = atomicCAS(&slot_pix[k], 0xFFFFFFFFu, pix);
                    if (v == 0xFFFFFFFFu || v == pix) {
                        slot = int(k);
                        break;
                    }
                }
                if (slot >= 0) {
                    double* h = hist + uint32_t(slot) * T3 + bin * 3u;
                    atomicAdd(h + 0, double(c.x));
                    atomicAdd(h + 1, double(c.y));
                    atomicAdd(h + 2, double(c.z));
                    pending = false;
                } else if (round == rounds - 1) {
                    film_add3(R.film, R.film_comp, (size_t(pix) * T + bin) * 3, c.x, c.y, c.z);
                    pending = false;
                }
            }
            __syncthreads();
            uint32_t used = 0;  // slots are claimed in order
            while (used < pix_slots && slot_pix[used] != 0xFFFFFFFFu) ++used;
            for (uint32_t i = threadIdx.x; i < used * T3; i += blockDim.x) {
                const double v = hist[i];
                if (v != 0.0) {
                    const uint32_t k = i / T3;
                    const float hi = float(v), lo = float(v - double(hi));
                    film_add_partial(R.film, R.film_comp, size_t(slot_pix[k]) * T3 + (i - k * T3), hi, lo);
                    hist[i] = 0.0;
                }
            }
            __syncthreads();
        }
    }
}

// multibounce films: visible connections mark their layer for the camera sample, add to the
// steady image, and (inside the time window) to their layer's film
__global__ void __launch_bounds__(kBlock) film_mb_kernel(RenderDev R, WaveDev Wv, int sh)
{
    const uint32_t count = Wv.counters[shword(sh)];
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count; r += gridDim.x * blockDim.x) {
        if (!Wv.sh_vis[r]) continue;
        const float4 c = Wv.sh_c[r];
        const uint32_t tag = Wv.sh_tag[r], pix = __float_as_uint(c.w), fidx = __float_as_uint(Wv.sh_d[r].w);
        atomicOr(R.cs_bits + (tag >> kMaxLayersTagBits), 1u << (tag & ((1u << kMaxLayersTagBits) - 1u)));
        if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
            film_add3(R.steady, R.steady_comp, size_t(pix) * 3, c.x, c.y, c.z);
            if (fidx != 0xFFFFFFFFu) {
                film_add3(R.film, R.film_comp, size_t(fidx) * 3, c.x, c.y, c.z);
            }
        }
    }
}

// multibounce films, after a wavefront: each camera sample adds 1 to the weight of every layer it
// sent samples to (a bounce film only sees those camera samples), then its marks are cleared
__global__ void __launch_bounds__(kBlock) mb_weight_kernel(RenderDev R, WaveDev Wv)
{
    const WaveCtl ctl = *Wv.ctl;
    const uint32_t n_cs = ctl.n / uint32_t(R.n_inc);
    const uint32_t cs0 = uint32_t(ctl.base / uint64_t(R.n_inc));
    const uint32_t pixels = R.layer_stride / uint32_t(R.T);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_cs; i += gridDim.x * blockDim.x) {
        uint32_t bits = R.cs_bits[i];
        if (!bits) continue;
        R.cs_bits[i] = 0u;
        const uint32_t pix = (cs0 + i) / uint32_t(R.spp2);
        for (int b = 0; b < R.layers; ++b)
            if (bits & (1u << b)) atomicAdd(R.layer_w + size_t(b) * pixels + pix, 1.0f);
    }
}

// batched renders: visible contributions go straight to their job's film, one compensated add each
__global__ void __launch_bounds__(kBlock) film_batch_kernel(const JobDev* __restrict__ jobs, WaveDev Wv, int sh)
{
    const uint32_t count = Wv.counters[shword(sh)];
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count; r += gridDim.x * blockDim.x) {
        if (!Wv.sh_vis[r]) continue;
        const RenderDev& R = jobs[Wv.sh_job[r]].R;
        float4 c = Wv.sh_c[r];
        uint32_t pix = __float_as_uint(c.w);
        uint32_t bin = __float_as_uint(Wv.sh_d[r].w) - pix * uint32_t(R.T);
        film_add3(R.film, R.film_comp, (size_t(pix) * uint32_t(R.T) + bin) * 3, c.x, c.y, c.z);
    }
}

// ---------------------------------------------------------------------------------------------
// film read-back: one block per pixel writes (sum + comp) / weight of its T x RGB bins (0/0 -> 0,
// ImageRow::weight, Image.h:153-162) in double precision, and the pixel's steady value (the sum
// over its bins) for the steady image (StreakCameraFilm.h:369-397)
constexpr int kNormBlock = 256;
__global__ void __launch_bounds__(kNormBlock) film_normalize_kernel(const float* __restrict__ sum,
                                                                   const float* __restrict__ comp,
                                                                   const double* __restrict__ weight, uint32_t T3,
                                                                   float* __restrict__ out, float* __restrict__ steady)
{
    __shared__ double part[kNormBlock / 64][3];
    const size_t p = blockIdx.x;
    const double w = weight[p];
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t j = threadIdx.x; j < T3; j += kNormBlock) {
        const size_t i = p * T3 + j;
        double v = (double(sum[i]) + double(comp[i])) / w;
        if (!isfinite(v)) v = 0.0;
        if (out) out[i] = float(v);
        const uint32_t k = j % 3u;
        acc[0] += k == 0 ? v : 0.0;
        acc[1] += k == 1 ? v : 0.0;
        acc[2] += k == 2 ? v : 0.0;
    }
    if (!steady) return;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off);
    if (__lane_id() == 0)
        for (int k = 0; k < 3; ++k) part[threadIdx.x >> 6][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int q = 0; q < kNormBlock / 64; ++q) t += part[q][threadIdx.x];
        steady[p * 3 + threadIdx.x] = float(t);
    }
}

// batch closest hit for the public query entry point
template <int TRAV, int STACK>
__global__ void __launch_bounds__(kTraceBlock) intersect_kernel(SceneDev S, const tr_ray* __restrict__ rays,
                                                                tr_hit* __restrict__ hits, const uint2* __restrict__ ids,
                                                                uint32_t n)
{
    __shared__ uint32_t stack[stack_words<TRAV, STACK>() * kTraceBlock];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        tr_ray q = rays[i];
        RayPre r = ray_pre(mk(q.org[0], q.org[1], q.org[2]), mk(q.dir[0], q.dir[1], q.dir[2]), S.plane_bound);
        float tfar = q.tfar;
        int32_t tri = trace_scene<false, TRAV>(S, r, q.tnear, tfar, stack);
        tr_hit h;
        if (tri >= 0) {
            uint2 gp = ids[tri];
            h.t = tfar;
            h.geom_id = gp.x;
            h.prim_id = gp.y;
            h.tri = uint32_t(tri);
        } else {
            h.t = -1.0f;
            h.geom_id = 0xFFFFFFFFu;
            h.prim_id = 0xFFFFFFFFu;
            h.tri = 0xFFFFFFFFu;
        }
        hits[i] = h;
    }
}

// any hit in (tnear, tfar] for each ray (the shadow kernel's query)
template <int TRAV, int STACK>
__global__ void __launch_bounds__(kTraceBlock) occluded_kernel(SceneDev S, const tr_ray* __restrict__ rays,
                                                               uint8_t* __restrict__ occluded, uint32_t n)
{
    __shared__ uint32_t stack[stack_words<TRAV, STACK>() * kTraceBlock];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        tr_ray q = rays[i];
        RayPre r = ray_pre(mk(q.org[0], q.org[1], q.org[2]), mk(q.dir[0], q.dir[1], q.dir[2]), S.plane_bound);
        float tfar = q.tfar;
        occluded[i] = trace_scene<true, TRAV>(S, r, q.tnear, tfar, stack) >= 0 ? 1 : 0;
    }
}

// The traversal stack holds at most one entry per BVH level; the smallest instantiation that
// covers the scene's depth keeps LDS per block (and so occupancy) in check.
// `bound` = entries the traversal stack can hold at most (binary: depth; 4-wide: 3 per level).
// Launches kernel k on `s`; with a start event, through hipExtLaunchKernelGGL, which stamps the
// start and stop events as part of the dispatch itself instead of as marker packets around it (a
// marker pair between two launches drains the GPU: 3 % of a media-workload render timed this way).
template <class... P, class... A>
void launch_timed(void (*k)(P...), uint32_t grid, uint32_t block, hipStream_t s, hipEvent_t e0, hipEvent_t e1, A... a)
{
    if (e0) hipExtLaunchKernelGGL(k, dim3(grid), dim3(block), 0u, s, e0, e1, 0u, P(a)...);
    else k<<<grid, block, 0, s>>>(a...);
}
// (the extend / shadow dispatch macros below run inside LaunchSet, whose ev_start / ev_stop are
// the current launch's events, or null)
#define TR_STACK_DISPATCH3(T, bound, B, KERNEL, GRID, BLOCK, STREAM, ...)                                   \
    do {                                                                                                    \
        if ((bound) < 16) launch_timed(KERNEL<T, 16, B>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__); \
        else if ((bound) < 24) launch_timed(KERNEL<T, 24, B>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__); \
        else if ((bound) < 32) launch_timed(KERNEL<T, 32, B>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__); \
        else launch_timed(KERNEL<T, kBvhStack, B>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__);      \
    } while (0)
#define TR_STACK_DISPATCH3B(T, bound, batch, KERNEL, GRID, BLOCK, STREAM, ...)                             \
    do {                                                                                                   \
        if (batch) TR_STACK_DISPATCH3(T, bound, true, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__);           \
        else TR_STACK_DISPATCH3(T, bound, false, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__);                \
    } while (0)
// extend / shadow launches: `trav` = kTrav* chosen on the host (trav_for)
#define TR_TRACE_DISPATCH(trav, bound, batch, KERNEL, GRID, BLOCK, STREAM, ...)                             \
    do {                                                                                                    \
        switch (trav) {                                                                                     \
        case kTravBinary: TR_STACK_DISPATCH3B(kTravBinary, bound, batch, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break; \
        case kTravWide: TR_STACK_DISPATCH3B(kTravWide, bound, batch, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break;     \
        case kTravWide16: TR_STACK_DISPATCH3B(kTravWide16, bound, batch, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break; \
        case kTravWidePP: TR_STACK_DISPATCH3B(kTravWidePP, bound, batch, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break; \
        default: TR_STACK_DISPATCH3B(kTravWideKeyed, bound, batch, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break;       \
        }                                                                                                   \
    } while (0)
// camera-ray packets (extend_camera_kernel), timed like the other extend launches
#define TR_CAMERA_DISPATCH_T(T, bound, GRID, BLOCK, STREAM, ...)                                                      \
    do {                                                                                                             \
        if ((bound) < 16) launch_timed(extend_camera_kernel<T, 16>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__); \
        else if ((bound) < 24) launch_timed(extend_camera_kernel<T, 24>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__); \
        else if ((bound) < 32) launch_timed(extend_camera_kernel<T, 32>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__); \
        else launch_timed(extend_camera_kernel<T, kBvhStack>, GRID, BLOCK, STREAM, ev_start, ev_stop, __VA_ARGS__);      \
    } while (0)
#define TR_STACK_DISPATCH_T(T, bound, KERNEL, GRID, BLOCK, STREAM, ...)                   \
    do {                                                                                   \
        if ((bound) < 16) KERNEL<T, 16><<<GRID, BLOCK, 0, STREAM>>>(__VA_ARGS__);          \
        else if ((bound) < 24) KERNEL<T, 24><<<GRID, BLOCK, 0, STREAM>>>(__VA_ARGS__);     \
        else if ((bound) < 32) KERNEL<T, 32><<<GRID, BLOCK, 0, STREAM>>>(__VA_ARGS__);     \
        else KERNEL<T, kBvhStack><<<GRID, BLOCK, 0, STREAM>>>(__VA_ARGS__);                \
    } while (0)
// public closest-hit queries: the scene's own traversal choice
#define TR_STACK_DISPATCH(trav, bound, KERNEL, GRID, BLOCK, STREAM, ...)                                  \
    do {                                                                                                  \
        switch (trav) {                                                                                   \
        case kTravBinary: TR_STACK_DISPATCH_T(kTravBinary, bound, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break; \
        case kTravWide: TR_STACK_DISPATCH_T(kTravWide, bound, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break;     \
        case kTravWide16: TR_STACK_DISPATCH_T(kTravWide16, bound, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break; \
        default: TR_STACK_DISPATCH_T(kTravWideKeyed, bound, KERNEL, GRID, BLOCK, STREAM, __VA_ARGS__); break;       \
        }                                                                                                 \
    } while (0)

}  // namespace

// =============================================================================================
// host side

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(TR_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

int check_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(TR_ERR_DEVICE, "no HIP device available");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(TR_ERR_DEVICE, std::string("device is ") + p.gcnArchName + ", the library is built for gfx950");
    return TR_OK;
}

template <class T>
int dev_alloc(T** p, size_t count, uint64_t* bytes = nullptr)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) return fail(TR_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    if (bytes) *bytes += count * sizeof(T);
    return TR_OK;
}

template <class T>
void dev_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Launch shape and scheduling choices of a render (the autotuner searches over these).
struct LaunchCfg {
    uint32_t wave = kDefaultWave;   // paths per wavefront
    uint32_t trace_per_cu = 24;     // persistent traversal blocks per compute unit (24: 1 % faster than 16 on 2^25-path wavefronts)
    uint32_t shade_per_cu = 2;      // shade blocks (1024 threads) per compute unit
    bool wide = true;               // 4-wide BVH
    bool graph = false;             // replay each wavefront's bounce loop from a captured hipGraph
    bool prof = false;              // per-kernel HIP events
    bool prof_extend = false;       // ... around the extend launches only
    int mode = kShadeRef;
    // 4-wide traversal stack entries: 0 = 16-bit when the BVH has at most 65,536 nodes (else
    // 32-bit); 1 = always 32-bit; 2 = keyed 16-bit for closest hits (16-bit for any hits)
    int stack_fmt = 0;
    // batches: render the jobs one after another (single-job kernels, LDS-privatised film) instead of
    // in one shared sequence of wavefronts
    bool per_job = false;
    // single-job renders: camera rays traced as per-pixel packets (extend_camera_kernel)
    bool camera_packets = true;
};

// launch flags -> stack format
static int stack_fmt_of(uint32_t flags)
{
    return (flags & TR_RENDER_POSTPONE_LEAVES) ? 3 : (flags & TR_RENDER_STACK32) ? 1 : (flags & TR_RENDER_STACK_KEYED) ? 2 : 0;
}
static uint32_t stack_flags_of(int fmt)
{
    return fmt == 1 ? TR_RENDER_STACK32 : fmt == 2 ? TR_RENDER_STACK_KEYED : fmt == 3 ? TR_RENDER_POSTPONE_LEAVES : 0u;
}

struct WaveBuffers {
    uint32_t cap = 0;
    float4 *ray_o[2] = {}, *ray_d[2] = {}, *thr[2] = {};
    uint4 *misc[2] = {}, *extra[2] = {};
    float2* hit = nullptr;
    float4 *sh_o = nullptr, *sh_d = nullptr, *sh_c = nullptr;
    uint8_t* sh_vis = nullptr;
    uint32_t* sh_tag = nullptr;  // allocated by the first guided render
    uint32_t* sh_job = nullptr;  // allocated by the first batched render
    WaveCtl* ctl = nullptr;
    JobDev* d_jobs = nullptr;
    // pinned host staging of the job table: the upload is asynchronous, so its source must stay
    // valid until the copy has run (every render ends with a stream synchronisation; a pageable
    // vector freed right after hipMemcpyAsync may still be read by the runtime's copy)
    JobDev* h_jobs = nullptr;
    uint32_t jobs_cap = 0;
    uint32_t* cs_bits = nullptr;  // multibounce films: layer marks per camera sample of a wavefront
    uint32_t cs_cap = 0;
    uint32_t* counters = nullptr;

    void release()
    {
        for (int b = 0; b < 2; ++b) {
            dev_free(ray_o[b]); dev_free(ray_d[b]); dev_free(thr[b]); dev_free(misc[b]); dev_free(extra[b]);
        }
        dev_free(hit); dev_free(sh_o); dev_free(sh_d); dev_free(sh_c); dev_free(sh_vis); dev_free(counters);
        dev_free(sh_tag); dev_free(sh_job); dev_free(ctl); dev_free(d_jobs); dev_free(cs_bits);
        if (h_jobs) (void)hipHostFree(h_jobs);
        h_jobs = nullptr;
        cap = 0;
        jobs_cap = 0;
        cs_cap = 0;
    }
    int ensure_cs(uint32_t n)
    {
        if (n <= cs_cap) return TR_OK;
        dev_free(cs_bits);
        int rc = dev_alloc(&cs_bits, n);
        if (rc) return rc;
        cs_cap = n;
        return TR_OK;
    }
    int ensure_batch(uint32_t n_jobs)
    {
        int rc;
        if (!sh_job && cap && (rc = dev_alloc(&sh_job, cap))) return rc;
        if (n_jobs + 1 > jobs_cap) {
            // an earlier upload from the pinned table may still be queued (a render that failed
            // before its final synchronisation): drain the device before freeing it
            if (h_jobs) {
                (void)hipDeviceSynchronize();
                (void)hipHostFree(h_jobs);
                h_jobs = nullptr;
            }
            dev_free(d_jobs);
            jobs_cap = 0;
            if ((rc = dev_alloc(&d_jobs, n_jobs + 1))) return rc;
            if (hipHostMalloc(reinterpret_cast<void**>(&h_jobs), size_t(n_jobs + 1) * sizeof(JobDev), hipHostMallocDefault) !=
                hipSuccess) {
                h_jobs = nullptr;
                dev_free(d_jobs);
                return fail(TR_ERR_NOMEM, "hipHostMalloc: job table");
            }
            jobs_cap = n_jobs + 1;
        }
        return TR_OK;
    }
    int ensure_tags()
    {
        if (sh_tag || cap == 0) return TR_OK;
        return dev_alloc(&sh_tag, cap);
    }
    int ensure(uint32_t n)
    {
        if (n <= cap) return TR_OK;
        release();
        int rc;
        for (int b = 0; b < 2; ++b)
            if ((rc = dev_alloc(&ray_o[b], n)) || (rc = dev_alloc(&ray_d[b], n)) || (rc = dev_alloc(&thr[b], n)) ||
                (rc = dev_alloc(&misc[b], n)) || (rc = dev_alloc(&extra[b], n))) {
                release();
                return rc;
            }
        if ((rc = dev_alloc(&hit, n)) || (rc = dev_alloc(&sh_o, n)) ||
            (rc = dev_alloc(&sh_d, n)) || (rc = dev_alloc(&sh_c, n)) || (rc = dev_alloc(&sh_vis, n)) ||
            (rc = dev_alloc(&counters, C_WORDS)) || (rc = dev_alloc(&ctl, 1))) {
            release();
            return rc;
        }
        cap = n;
        return TR_OK;
    }
    WaveDev view() const
    {
        WaveDev w;
        for (int b = 0; b < 2; ++b) {
            w.ray_o[b] = ray_o[b]; w.ray_d[b] = ray_d[b]; w.thr[b] = thr[b]; w.misc[b] = misc[b]; w.extra[b] = extra[b];
        }
        w.hit = hit;
        w.sh_o = sh_o; w.sh_d = sh_d; w.sh_c = sh_c; w.sh_vis = sh_vis; w.sh_tag = sh_tag;
        w.sh_job = sh_job; w.ctl = ctl;
        w.counters = counters;
        return w;
    }
};

}  // namespace

namespace tri {
int fail(int code, const std::string& msg) { return ::fail(code, msg); }
int check_device() { return ::check_device(); }
int compute_units()
{
    int dev = 0, cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) return 256;
    return cu;
}
}  // namespace tri

struct tr_scene_s {
    std::vector<float> tris;        // n*9
    std::vector<int32_t> tri_mat;   // material index | kMatTwoSided | kMatSmooth
    std::vector<uint32_t> tri_geom, tri_prim;
    std::vector<float> tri_vn;      // per triangle corner shading normals (9 floats), once a mesh has some
    std::vector<float> mats;        // n_mat*4
    std::vector<LightHost> lights;
    uint32_t n_geoms = 0;
    bool has_medium = false;
    float sa[3] = {0, 0, 0}, ss[3] = {0, 0, 0};
    bool committed = false;
    hipStream_t stream = nullptr;
    float4 *d_nodes = nullptr, *d_nodes4 = nullptr, *d_tris = nullptr, *d_tri_normal = nullptr, *d_mats = nullptr;
    float4* d_tri_vn = nullptr;  // smooth triangles: p0, e1, e2, n0, n1, n2 (6 float4 per triangle)
    LightHost* d_lights = nullptr;
    uint2* d_ids = nullptr;
    uint32_t n_nodes = 0, n_nodes4 = 0;
    int depth = 0, depth4 = 0, stack4 = 0;
    bool wide = true;   // traverse the 4-wide BVH (tr_render_desc flags can select the binary one)
    float bb_lo[3] = {0, 0, 0}, bb_hi[3] = {1, 1, 1};  // bounds of the triangles (set at commit)
    float plane_bound = 1.0f;  // largest |plane| of the quantized BVH (set at commit)
    LaunchCfg tune;     // launch shape used by renders of this scene (tr_autotune sets it)
    double build_ms = 0.0;
    uint64_t dev_bytes = 0;
    WaveBuffers wave;
    hipEvent_t ev[2 * 256] = {};
    bool ev_made = false;
    int compute_units = 256;

    void release_device()
    {
        dev_free(d_nodes); dev_free(d_nodes4); dev_free(d_tris); dev_free(d_tri_normal); dev_free(d_mats); dev_free(d_lights);
        dev_free(d_tri_vn);
        dev_free(d_ids);
        dev_bytes = 0;
    }
    SceneDev view() const
    {
        SceneDev s;
        s.nodes = d_nodes;
        s.nodes4 = d_nodes4;
        s.tris = d_tris;
        s.tri_normal = d_tri_normal;
        s.tri_vn = d_tri_vn;
        s.mats = d_mats;
        s.lights = d_lights;
        s.n_lights = int32_t(lights.size());
        s.has_medium = has_medium ? 1 : 0;
        s.sigma_s = make_float3(ss[0], ss[1], ss[2]);
        s.sigma_t = make_float3(sa[0] + ss[0], sa[1] + ss[1], sa[2] + ss[2]);
        s.lambda = (s.sigma_t.x + s.sigma_t.y + s.sigma_t.z) / 3.0f;
        s.medium_albedo = has_medium ? (ss[0] / s.sigma_t.x + ss[1] / s.sigma_t.y + ss[2] / s.sigma_t.z) / 3.0f : 0.f;
        s.plane_bound = plane_bound;
        return s;
    }
};

struct tr_film_s {
    tr_film_desc desc{};
    int rows = 0;
    int layers = 0;      // multibounce bounce films (0: one streak film)
    size_t n_elems = 0;  // max(layers, 1) * rows*W*T*3
    float *d_sum = nullptr, *d_comp = nullptr;        // compensated sums: sum + comp = film
    float *d_layer_w = nullptr, *d_steady = nullptr;  // multibounce films only (steady: sums, then comps)
    float *d_read = nullptr, *d_read_steady = nullptr;  // read-back staging (allocated by the first read)
    size_t read_cap = 0;                                // floats d_read holds
    double* d_read_w = nullptr;
    std::vector<double> weight;  // rows*W
    hipStream_t stream = nullptr;
};

extern "C" {

const char* tr_last_error(void) { return g_err.c_str(); }

int tr_device_info(char* name, int name_len, int* compute_units, uint64_t* global_mem)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(TR_ERR_DEVICE, "no HIP device available");
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, 0));
    if (name && name_len > 0) {
        std::snprintf(name, size_t(name_len), "%s (%s)", p.name, p.gcnArchName);
    }
    if (compute_units) *compute_units = p.multiProcessorCount;
    if (global_mem) *global_mem = p.totalGlobalMem;
    return TR_OK;
}

int tr_scene_create(tr_scene* out)
{
    if (!out) return fail(TR_ERR_INVALID, "tr_scene_create: null output");
    *out = nullptr;
    int rc = check_device();
    if (rc) return rc;
    tr_scene s = new (std::nothrow) tr_scene_s();
    if (!s) return fail(TR_ERR_NOMEM, "tr_scene_create: host allocation");
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(TR_ERR_DEVICE, "tr_scene_create: hipStreamCreate failed");
    }
    {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            s->compute_units = p.multiProcessorCount;
    }
    *out = s;
    return TR_OK;
}

int tr_scene_destroy(tr_scene s)
{
    if (!s) return TR_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->release_device();
    s->wave.release();
    if (s->ev_made)
        for (auto& e : s->ev) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return TR_OK;
}

int tr_scene_add_mesh(tr_scene s, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const tr_material* m, uint32_t* geom_id)
{
    return tr_scene_add_mesh_ex(s, xyz, nv, idx, nt, nullptr, m, 0u, geom_id);
}

int tr_scene_add_obj(tr_scene s, const char* path, const tr_material* m, uint32_t flags, uint32_t* geom_id)
{
    if (!s || !path || !m) return fail(TR_ERR_INVALID, "tr_scene_add_obj: null argument");
    tri::ObjMesh mesh;
    int rc = tri::load_obj(path, &mesh);
    if (rc) return rc;
    return tr_scene_add_mesh_ex(s, mesh.xyz.data(), uint32_t(mesh.xyz.size() / 3), mesh.idx.data(),
                                uint32_t(mesh.idx.size() / 3), mesh.has_normals ? mesh.normals.data() : nullptr, m,
                                flags, geom_id);
}

int tr_scene_add_mesh_ex(tr_scene s, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt,
                         const float* normals, const tr_material* m, uint32_t flags, uint32_t* geom_id)
{
    if (!s || (nt > 0 && (!xyz || !idx)) || !m) return fail(TR_ERR_INVALID, "tr_scene_add_mesh: null argument");
    if (flags & ~uint32_t(TR_MATERIAL_TWO_SIDED)) return fail(TR_ERR_INVALID, "tr_scene_add_mesh: unknown flags");
    for (uint32_t i = 0; i < 3 * nt; ++i)
        if (idx[i] >= nv) return fail(TR_ERR_INVALID, "tr_scene_add_mesh: vertex index out of range");
    for (int k = 0; k < 3; ++k)
        if (!(m->kd[k] >= 0.f && m->kd[k] <= 1.f)) return fail(TR_ERR_INVALID, "tr_scene_add_mesh: Kd outside [0,1]");
    if (s->tris.size() / 9 + nt > 0x7FFFFFFFu) return fail(TR_ERR_INVALID, "tr_scene_add_mesh: too many triangles");
    int32_t mat = int32_t(s->mats.size() / 4);
    // mean albedo used by the roulette: mean of 1 - (1 - Kd) (BidirectionalPathTracing.h:632-638)
    float alb = ((1.f - (1.f - m->kd[0])) + (1.f - (1.f - m->kd[1])) + (1.f - (1.f - m->kd[2]))) / 3.0f;
    s->mats.insert(s->mats.end(), {m->kd[0], m->kd[1], m->kd[2], alb});
    uint32_t g = s->n_geoms++;
    // material index with the two-sided flag (bit 30) and, per triangle, the smooth flag (bit 29)
    const int32_t mat_bits = mat | ((flags & TR_MATERIAL_TWO_SIDED) ? kMatTwoSided : 0);
    if (normals && s->tri_vn.empty()) s->tri_vn.assign(s->tri_mat.size() * 9, 0.f);
    for (uint32_t t = 0; t < nt; ++t) {
        for (int k = 0; k < 3; ++k) {
            const float* p = xyz + 3 * size_t(idx[3 * t + k]);
            s->tris.insert(s->tris.end(), {p[0], p[1], p[2]});
        }
        bool smooth = false;  // all-zero corner normals mean a flat triangle
        if (normals)
            for (int k = 0; k < 9; ++k) smooth = smooth || normals[9 * size_t(t) + k] != 0.f;
        s->tri_mat.push_back(smooth ? (mat_bits | kMatSmooth) : mat_bits);
        s->tri_geom.push_back(g);
        s->tri_prim.push_back(t);
        if (!s->tri_vn.empty()) {
            if (normals) s->tri_vn.insert(s->tri_vn.end(), normals + 9 * size_t(t), normals + 9 * size_t(t) + 9);
            else s->tri_vn.insert(s->tri_vn.end(), 9, 0.f);
        }
    }
    if (geom_id) *geom_id = g;
    s->committed = false;
    return TR_OK;
}

int tr_scene_add_point_light(tr_scene s, const float position[3], const float intensity[3])
{
    if (!s || !position || !intensity) return fail(TR_ERR_INVALID, "tr_scene_add_point_light: null argument");
    LightHost L{};
    std::memcpy(L.pos, position, 12);
    std::memcpy(L.I, intensity, 12);
    L.I[3] = float(kLightPoint);
    s->lights.push_back(L);
    s->committed = false;
    return TR_OK;
}

int tr_scene_add_rect_light(tr_scene s, const float corner[3], const float target[3], const float size[2],
                            float intensity)
{
    if (!s || !corner || !target || !size) return fail(TR_ERR_INVALID, "tr_scene_add_rect_light: null argument");
    if (!(size[0] > 0.f && size[1] > 0.f)) return fail(TR_ERR_INVALID, "tr_scene_add_rect_light: non-positive size");
    // RectangularAreaLightSource::setup (RectangularAreaLightSource.h:66-80)
    double n[3] = {double(target[0]) - corner[0], double(target[1]) - corner[1], double(target[2]) - corner[2]};
    float nf[3] = {float(target[0] - corner[0]), float(target[1] - corner[1]), float(target[2] - corner[2])};
    float len = std::sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
    if (!(len > 0.f)) return fail(TR_ERR_INVALID, "tr_scene_add_rect_light: target equals corner");
    (void)n;
    float inv = float(1. / double(len));
    float N[3] = {nf[0] * inv, nf[1] * inv, nf[2] * inv};
    float up[3] = {0.f, 1.f, 0.f};
    if (!(std::fabs(double(N[1])) < 0.9999)) {
        up[1] = 0.f;
        up[2] = 1.f;
    }
    auto crs = [](const float* a, const float* b, float* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    float left[3], up2[3];
    crs(N, up, left);
    crs(N, left, up2);
    LightHost L{};
    std::memcpy(L.pos, corner, 12);
    for (int k = 0; k < 3; ++k) {
        L.left[k] = left[k] * size[0];
        L.up[k] = up2[k] * size[1];
        L.normal[k] = N[k];
        L.I[k] = intensity;
    }
    L.normal[3] = float(1. / double(size[0] * size[1]));
    L.I[3] = float(kLightRect);
    s->lights.push_back(L);
    s->committed = false;
    return TR_OK;
}

int tr_scene_set_medium(tr_scene s, const float sigma_a[3], const float sigma_s[3])
{
    if (!s) return fail(TR_ERR_INVALID, "tr_scene_set_medium: null scene");
    if (!sigma_a || !sigma_s) {
        s->has_medium = false;
        return TR_OK;
    }
    for (int k = 0; k < 3; ++k)
        if (!(sigma_a[k] >= 0.f && sigma_s[k] >= 0.f && sigma_a[k] + sigma_s[k] > 0.f))
            return fail(TR_ERR_INVALID, "tr_scene_set_medium: coefficients must be >= 0 with positive extinction");
    std::memcpy(s->sa, sigma_a, 12);
    std::memcpy(s->ss, sigma_s, 12);
    s->has_medium = true;
    s->committed = false;
    return TR_OK;
}

int tr_scene_commit(tr_scene s)
{
    if (!s) return fail(TR_ERR_INVALID, "tr_scene_commit: null scene");
    uint32_t n = uint32_t(s->tri_mat.size());
    BvhBuildResult b = build_bvh(s->tris.data(), s->tri_mat.data(), n);
    if (b.depth >= kBvhStack) return fail(TR_ERR_INVALID, "tr_scene_commit: BVH deeper than the traversal stack");
    Bvh4 b4 = collapse_bvh4(b);
    if (b4.stack_bound >= kBvhStack) return fail(TR_ERR_INVALID, "tr_scene_commit: 4-wide BVH deeper than the traversal stack");
    if (b4.nodes.size() >= (size_t(1) << 29) || b.tris.size() >= (size_t(1) << 29))
        return fail(TR_ERR_INVALID, "tr_scene_commit: more than 2^29 nodes or triangles");
    const std::vector<Bvh4QNodeHost> q4 = quantize_bvh4(b4);
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->release_device();
    std::vector<float4> normals(n > 0 ? n : 1);
    for (uint32_t i = 0; i < n; ++i) {
        const float* t = &s->tris[9 * size_t(i)];
        // ng = normalize(cross(p2 - p0, p1 - p0)) (World.h:449)
        float e2[3] = {t[6] - t[0], t[7] - t[1], t[8] - t[2]}, e1[3] = {t[3] - t[0], t[4] - t[1], t[5] - t[2]};
        float c[3] = {e2[1] * e1[2] - e2[2] * e1[1], e2[2] * e1[0] - e2[0] * e1[2], e2[0] * e1[1] - e2[1] * e1[0]};
        float l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        float inv = l > 0.f ? float(1. / double(l)) : 0.f;
        int32_t mi = s->tri_mat[i];
        float w;
        std::memcpy(&w, &mi, 4);
        normals[i] = make_float4(c[0] * inv, c[1] * inv, c[2] * inv, w);
    }
    for (int k = 0; k < 3; ++k) {
        s->bb_lo[k] = n ? 3.4e38f : 0.0f;
        s->bb_hi[k] = n ? -3.4e38f : 1.0f;
    }
    for (size_t v = 0; v < size_t(n) * 3; ++v)
        for (int k = 0; k < 3; ++k) {
            s->bb_lo[k] = std::min(s->bb_lo[k], s->tris[3 * v + size_t(k)]);
            s->bb_hi[k] = std::max(s->bb_hi[k], s->tris[3 * v + size_t(k)]);
        }
    std::vector<uint2> ids(n > 0 ? n : 1);
    for (uint32_t i = 0; i < n; ++i) ids[i] = make_uint2(s->tri_geom[i], s->tri_prim[i]);
    size_t nn = b.nodes.size(), nt = b.tris.size(), nm = std::max<size_t>(1, s->mats.size() / 4),
           nl = std::max<size_t>(1, s->lights.size());
    int rc;
    if ((rc = dev_alloc(&s->d_nodes, 4 * nn, &s->dev_bytes)) || (rc = dev_alloc(&s->d_nodes4, 4 * q4.size(), &s->dev_bytes)) || (rc = dev_alloc(&s->d_tris, 4 * std::max<size_t>(1, nt), &s->dev_bytes)) ||
        (rc = dev_alloc(&s->d_tri_normal, normals.size(), &s->dev_bytes)) || (rc = dev_alloc(&s->d_mats, nm, &s->dev_bytes)) ||
        (rc = dev_alloc(&s->d_lights, nl, &s->dev_bytes)) || (rc = dev_alloc(&s->d_ids, ids.size(), &s->dev_bytes))) {
        s->release_device();
        return rc;
    }
    HIP_TRY(hipMemcpyAsync(s->d_nodes, b.nodes.data(), nn * sizeof(BvhNodeHost), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_nodes4, q4.data(), q4.size() * sizeof(Bvh4QNodeHost), hipMemcpyHostToDevice, s->stream));
    // rotated triangle records (TriRot): per vertex x y z x y, then the scene triangle index
    std::vector<float> trot(16 * std::max<size_t>(1, nt), 0.f);
    for (size_t i = 0; i < nt; ++i) {
        const TriHost& t = b.tris[i];
        const float* P[3] = {t.p0, t.p1, t.p2};
        float* o = &trot[16 * i];
        for (int v = 0; v < 3; ++v) {
            o[5 * v + 0] = P[v][0];
            o[5 * v + 1] = P[v][1];
            o[5 * v + 2] = P[v][2];
            o[5 * v + 3] = P[v][0];
            o[5 * v + 4] = P[v][1];
        }
        o[15] = t.p0[3];  // scene triangle index bits
    }
    if (nt) HIP_TRY(hipMemcpyAsync(s->d_tris, trot.data(), nt * 16 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_tri_normal, normals.data(), normals.size() * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    if (!s->mats.empty())
        HIP_TRY(hipMemcpyAsync(s->d_mats, s->mats.data(), s->mats.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    if (!s->lights.empty())
        HIP_TRY(hipMemcpyAsync(s->d_lights, s->lights.data(), s->lights.size() * sizeof(LightHost), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->d_ids, ids.data(), ids.size() * sizeof(uint2), hipMemcpyHostToDevice, s->stream));
    std::vector<float4> vn;
    if (!s->tri_vn.empty()) {
        // smooth triangles: the plane frame for barycentrics from a hit point, and corner normals
        vn.resize(size_t(n) * 6);
        for (uint32_t i = 0; i < n; ++i) {
            const float* t = &s->tris[9 * size_t(i)];
            const float* q = &s->tri_vn[9 * size_t(i)];
            vn[6 * size_t(i) + 0] = make_float4(t[0], t[1], t[2], 0.f);
            vn[6 * size_t(i) + 1] = make_float4(t[3] - t[0], t[4] - t[1], t[5] - t[2], 0.f);
            vn[6 * size_t(i) + 2] = make_float4(t[6] - t[0], t[7] - t[1], t[8] - t[2], 0.f);
            for (int k = 0; k < 3; ++k) vn[6 * size_t(i) + 3 + k] = make_float4(q[3 * k], q[3 * k + 1], q[3 * k + 2], 0.f);
        }
        if ((rc = dev_alloc(&s->d_tri_vn, vn.size(), &s->dev_bytes))) {
            s->release_device();
            return rc;
        }
        HIP_TRY(hipMemcpyAsync(s->d_tri_vn, vn.data(), vn.size() * sizeof(float4), hipMemcpyHostToDevice, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->n_nodes = uint32_t(nn);
    s->depth = b.depth;
    s->n_nodes4 = uint32_t(b4.nodes.size());
    s->depth4 = b4.depth;
    s->stack4 = b4.stack_bound;
    s->plane_bound = qbvh_plane_bound(q4);
    s->build_ms = b.build_ms;
    s->committed = true;
    return TR_OK;
}

int tr_scene_info_get(tr_scene s, tr_scene_info* info)
{
    if (!s || !info) return fail(TR_ERR_INVALID, "tr_scene_info_get: null argument");
    info->n_tris = uint32_t(s->tri_mat.size());
    info->n_bvh_nodes = s->n_nodes;
    info->bvh_depth = uint32_t(s->depth);
    info->n_lights = uint32_t(s->lights.size());
    info->device_bytes = s->dev_bytes;
    info->build_ms = s->build_ms;
    return TR_OK;
}

static uint32_t trace_grid(uint32_t n)
{
    uint32_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    return std::max(1u, std::min(blocks, 256u * 40u));
}

// Persistent traversal grids: enough blocks to fill every CU (each wavefront then keeps fetching
// queue entries), never more blocks than the wavefront has rays.
static uint32_t persistent_grid(uint32_t n, int compute_units)
{
    uint32_t blocks = (n + kTraceBlock - 1) / kTraceBlock;
    return std::max(1u, std::min(blocks, uint32_t(std::max(compute_units, 1)) * 16u));
}

int tr_intersect_batch_device(tr_scene s, const tr_ray* d_rays, uint32_t n, tr_hit* d_hits, void* stream)
{
    if (!s) return fail(TR_ERR_INVALID, "tr_intersect_batch_device: null scene");
    if (!s->committed) return fail(TR_ERR_STATE, "tr_intersect_batch_device: scene not committed");
    if (n == 0) return TR_OK;
    if (!d_rays || !d_hits) return fail(TR_ERR_INVALID, "tr_intersect_batch_device: null buffer");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    const int trav = !s->wide ? kTravBinary : s->n_nodes4 <= 65536u ? kTravWide16 : kTravWide;
    TR_STACK_DISPATCH(trav, s->wide ? s->stack4 : s->depth, intersect_kernel, trace_grid(n), kTraceBlock, st, s->view(),
                      d_rays, d_hits, s->d_ids, n);
    HIP_TRY(hipGetLastError());
    return TR_OK;
}

int tr_intersect_batch(tr_scene s, const tr_ray* rays, uint32_t n, tr_hit* hits)
{
    if (!s) return fail(TR_ERR_INVALID, "tr_intersect_batch: null scene");
    if (!s->committed) return fail(TR_ERR_STATE, "tr_intersect_batch: scene not committed");
    if (n == 0) return TR_OK;
    if (!rays || !hits) return fail(TR_ERR_INVALID, "tr_intersect_batch: null buffer");
    tr_ray* dr = nullptr;
    tr_hit* dh = nullptr;
    int rc;
    if ((rc = dev_alloc(&dr, n)) || (rc = dev_alloc(&dh, n))) {
        dev_free(dr);
        return rc;
    }
    int out = TR_OK;
    if (hipMemcpyAsync(dr, rays, size_t(n) * sizeof(tr_ray), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        out = fail(TR_ERR_DEVICE, "tr_intersect_batch: upload failed");
    if (!out) out = tr_intersect_batch_device(s, dr, n, dh, s->stream);
    if (!out && hipMemcpyAsync(hits, dh, size_t(n) * sizeof(tr_hit), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        out = fail(TR_ERR_DEVICE, "tr_intersect_batch: download failed");
    if (!out && hipStreamSynchronize(s->stream) != hipSuccess) out = fail(TR_ERR_DEVICE, "tr_intersect_batch: kernel failed");
    dev_free(dr);
    dev_free(dh);
    return out;
}

}  // extern "C"

// ---- closest-hit queries without allocation ---------------------------------------------------
// Staging for tr_intersect_batch_noalloc: rays, then hits, of at most TR_NOALLOC_MAX_RAYS rays, in
// one static device array per GPU (held until the process ends). Host code never passes the
// array's name to a kernel or a copy: it takes the device address from hipGetSymbolAddress once
// per device and uses that.
namespace {
constexpr size_t kNoAllocRayBytes = size_t(TR_NOALLOC_MAX_RAYS) * sizeof(tr_ray);
constexpr size_t kNoAllocBytes = kNoAllocRayBytes + size_t(TR_NOALLOC_MAX_RAYS) * sizeof(tr_hit);
static_assert(kNoAllocRayBytes % 256 == 0, "hits start 256-byte aligned");
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch array (looked up on the first call on each device)
int scratch_on_current_device(void** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TR_ERR_DEVICE, "tr_intersect_batch_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(TR_ERR_DEVICE, std::string("tr_intersect_batch_noalloc: hipGetSymbolAddress: ") +
                                           hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TR_OK;
}
}  // namespace

extern "C" {

int tr_intersect_batch_noalloc(tr_scene s, const tr_ray* rays, uint32_t n, tr_hit* hits)
{
    if (!s) return fail(TR_ERR_INVALID, "tr_intersect_batch_noalloc: null scene");
    if (!s->committed) return fail(TR_ERR_STATE, "tr_intersect_batch_noalloc: scene not committed");
    if (n == 0) return TR_OK;
    if (!rays || !hits) return fail(TR_ERR_INVALID, "tr_intersect_batch_noalloc: null buffer");
    if (n > TR_NOALLOC_MAX_RAYS)
        return fail(TR_ERR_INVALID, "tr_intersect_batch_noalloc: " + std::to_string(n) + " rays exceed the limit of " +
                                        std::to_string(TR_NOALLOC_MAX_RAYS) + " per call");
    void* base = nullptr;
    int rc = scratch_on_current_device(&base);
    if (rc) return rc;
    tr_ray* dr = static_cast<tr_ray*>(base);
    tr_hit* dh = reinterpret_cast<tr_hit*>(static_cast<unsigned char*>(base) + kNoAllocRayBytes);
    int out = TR_OK;
    if (hipMemcpyAsync(dr, rays, size_t(n) * sizeof(tr_ray), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        out = fail(TR_ERR_DEVICE, "tr_intersect_batch_noalloc: upload failed");
    if (!out) out = tr_intersect_batch_device(s, dr, n, dh, s->stream);
    if (!out && hipMemcpyAsync(hits, dh, size_t(n) * sizeof(tr_hit), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        out = fail(TR_ERR_DEVICE, "tr_intersect_batch_noalloc: download failed");
    // also on failure: the queued copies read and write the caller's buffers
    if (hipStreamSynchronize(s->stream) != hipSuccess && !out)
        out = fail(TR_ERR_DEVICE, "tr_intersect_batch_noalloc: kernel failed");
    return out;
}

int tr_occluded_batch_device(tr_scene s, const tr_ray* d_rays, uint32_t n, uint8_t* d_occ, void* stream)
{
    if (!s) return fail(TR_ERR_INVALID, "tr_occluded_batch_device: null scene");
    if (!s->committed) return fail(TR_ERR_STATE, "tr_occluded_batch_device: scene not committed");
    if (n == 0) return TR_OK;
    if (!d_rays || !d_occ) return fail(TR_ERR_INVALID, "tr_occluded_batch_device: null buffer");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    const int trav = !s->wide ? kTravBinary : s->n_nodes4 <= 65536u ? kTravWide16 : kTravWide;
    TR_STACK_DISPATCH(trav, s->wide ? s->stack4 : s->depth, occluded_kernel, trace_grid(n), kTraceBlock, st, s->view(),
                      d_rays, d_occ, n);
    HIP_TRY(hipGetLastError());
    return TR_OK;
}

int tr_occluded_batch(tr_scene s, const tr_ray* rays, uint32_t n, uint8_t* occluded)
{
    if (!s) return fail(TR_ERR_INVALID, "tr_occluded_batch: null scene");
    if (!s->committed) return fail(TR_ERR_STATE, "tr_occluded_batch: scene not committed");
    if (n == 0) return TR_OK;
    if (!rays || !occluded) return fail(TR_ERR_INVALID, "tr_occluded_batch: null buffer");
    tr_ray* dr = nullptr;
    uint8_t* dq = nullptr;
    int rc;
    if ((rc = dev_alloc(&dr, n)) || (rc = dev_alloc(&dq, n))) {
        dev_free(dr);
        return rc;
    }
    int out = TR_OK;
    if (hipMemcpyAsync(dr, rays, size_t(n) * sizeof(tr_ray), hipMemcpyHostToDevice, s->stream) != hipSuccess)
        out = fail(TR_ERR_DEVICE, "tr_occluded_batch: upload failed");
    if (!out) out = tr_occluded_batch_device(s, dr, n, dq, s->stream);
    if (!out && hipMemcpyAsync(occluded, dq, size_t(n), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
        out = fail(TR_ERR_DEVICE, "tr_occluded_batch: download failed");
    if (!out && hipStreamSynchronize(s->stream) != hipSuccess) out = fail(TR_ERR_DEVICE, "tr_occluded_batch: kernel failed");
    dev_free(dr);
    dev_free(dq);
    return out;
}

int tr_film_create(tr_film* out, const tr_film_desc* d)
{
    if (!out || !d) return fail(TR_ERR_INVALID, "tr_film_create: null argument");
    *out = nullptr;
    if (d->width <= 0 || d->height <= 0 || d->time_bins <= 0 || !(d->exposure > 0.f) ||
        d->scanline >= d->height || d->scanline < -1)
        return fail(TR_ERR_INVALID, "tr_film_create: bad film geometry");
    if (d->bounce_layers < 0 || d->bounce_layers > 32)
        return fail(TR_ERR_INVALID, "tr_film_create: bounce_layers must be in [0, 32]");
    int rc = check_device();
    if (rc) return rc;
    int rows = d->scanline >= 0 ? 1 : d->height;
    const int layers = d->bounce_layers;
    size_t n = size_t(std::max(layers, 1)) * size_t(rows) * size_t(d->width) * size_t(d->time_bins) * 3;
    if (n / 3 >= 0xFFFFFFFFull) return fail(TR_ERR_INVALID, "tr_film_create: film larger than 2^32 bins");
    tr_film f = new (std::nothrow) tr_film_s();
    if (!f) return fail(TR_ERR_NOMEM, "tr_film_create: host allocation");
    f->desc = *d;
    f->rows = rows;
    f->layers = layers;
    f->n_elems = n;
    auto release = [&]() {
        dev_free(f->d_sum); dev_free(f->d_comp); dev_free(f->d_layer_w); dev_free(f->d_steady);
    };
    if ((rc = dev_alloc(&f->d_sum, n)) || (rc = dev_alloc(&f->d_comp, n)) ||
        (layers && ((rc = dev_alloc(&f->d_layer_w, size_t(layers) * rows * size_t(d->width))) ||
                    (rc = dev_alloc(&f->d_steady, size_t(rows) * size_t(d->width) * 6))))) {
        release();
        delete f;
        return rc;
    }
    if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) {
        release();
        delete f;
        return fail(TR_ERR_DEVICE, "tr_film_create: hipStreamCreate failed");
    }
    f->weight.assign(size_t(rows) * size_t(d->width), 0.0);
    *out = f;
    rc = tr_film_clear(f);
    if (rc) {
        tr_film_destroy(f);
        *out = nullptr;
    }
    return rc;
}

int tr_film_destroy(tr_film f)
{
    if (!f) return TR_OK;
    if (f->stream) (void)hipStreamSynchronize(f->stream);
    dev_free(f->d_sum); dev_free(f->d_comp); dev_free(f->d_layer_w); dev_free(f->d_steady);
    dev_free(f->d_read); dev_free(f->d_read_steady); dev_free(f->d_read_w);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
    return TR_OK;
}

int tr_film_clear(tr_film f)
{
    if (!f) return fail(TR_ERR_INVALID, "tr_film_clear: null film");
    const size_t bytes = f->n_elems * sizeof(float);
    HIP_TRY(hipMemsetAsync(f->d_sum, 0, bytes, f->stream));
    HIP_TRY(hipMemsetAsync(f->d_comp, 0, bytes, f->stream));
    if (f->layers) {
        const size_t px = size_t(f->rows) * size_t(f->desc.width);
        HIP_TRY(hipMemsetAsync(f->d_layer_w, 0, size_t(f->layers) * px * sizeof(float), f->stream));
        HIP_TRY(hipMemsetAsync(f->d_steady, 0, px * 6 * sizeof(float), f->stream));
    }
    HIP_TRY(hipStreamSynchronize(f->stream));
    std::fill(f->weight.begin(), f->weight.end(), 0.0);
    return TR_OK;
}

int tr_film_read_sums(tr_film f, double* sums, double* weights)
{
    if (!f) return fail(TR_ERR_INVALID, "tr_film_read_sums: null film");
    if (sums) {
        std::vector<float> s(f->n_elems), c(f->n_elems);
        HIP_TRY(hipMemcpyAsync(s.data(), f->d_sum, f->n_elems * sizeof(float), hipMemcpyDeviceToHost, f->stream));
        HIP_TRY(hipMemcpyAsync(c.data(), f->d_comp, f->n_elems * sizeof(float), hipMemcpyDeviceToHost, f->stream));
        HIP_TRY(hipStreamSynchronize(f->stream));
        for (size_t i = 0; i < f->n_elems; ++i) sums[i] = double(s[i]) + double(c[i]);
    }
    if (weights) {
        if (f->layers) {
            const size_t n = size_t(f->layers) * f->weight.size();
            std::vector<float> lw(n);
            HIP_TRY(hipMemcpyAsync(lw.data(), f->d_layer_w, n * sizeof(float), hipMemcpyDeviceToHost, f->stream));
            HIP_TRY(hipStreamSynchronize(f->stream));
            for (size_t i = 0; i < n; ++i) weights[i] = double(lw[i]);
        } else {
            std::copy(f->weight.begin(), f->weight.end(), weights);
        }
    }
    return TR_OK;
}

int tr_film_read_rows(tr_film f, int32_t row_begin, int32_t row_count, float* out, float* steady)
{
    if (!f) return fail(TR_ERR_INVALID, "tr_film_read_rows: null film");
    if (f->layers) return fail(TR_ERR_INVALID, "tr_film_read_rows: multibounce films are read whole (tr_film_read)");
    if (row_begin < 0 || row_count < 0 || row_begin + int64_t(row_count) > f->rows)
        return fail(TR_ERR_INVALID, "tr_film_read_rows: rows outside the film");
    if ((!out && !steady) || row_count == 0) return TR_OK;
    const size_t W = size_t(f->desc.width), T = size_t(f->desc.time_bins), P = size_t(f->rows) * W;
    const size_t p0 = size_t(row_begin) * W, np = size_t(row_count) * W, n = np * T * 3;
    // normalised on the device; one copy of the result instead of both raw arrays
    int rc;
    if (!f->d_read_w && (rc = dev_alloc(&f->d_read_w, P))) return rc;
    if (out && f->read_cap < n) {
        dev_free(f->d_read);
        f->read_cap = 0;
        if ((rc = dev_alloc(&f->d_read, n))) return rc;
        f->read_cap = n;
    }
    if (steady && !f->d_read_steady && (rc = dev_alloc(&f->d_read_steady, P * 3))) return rc;
    HIP_TRY(hipMemcpyAsync(f->d_read_w, f->weight.data(), P * sizeof(double), hipMemcpyHostToDevice, f->stream));
    film_normalize_kernel<<<uint32_t(np), kNormBlock, 0, f->stream>>>(f->d_sum + p0 * T * 3, f->d_comp + p0 * T * 3,
                                                                     f->d_read_w + p0, uint32_t(3 * T),
                                                                     out ? f->d_read : nullptr,
                                                                     steady ? f->d_read_steady : nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && out) e = hipMemcpyAsync(out, f->d_read, n * sizeof(float), hipMemcpyDeviceToHost, f->stream);
    if (e == hipSuccess && steady)
        e = hipMemcpyAsync(steady, f->d_read_steady, np * 3 * sizeof(float), hipMemcpyDeviceToHost, f->stream);
    // the weight upload reads f->weight asynchronously: always wait
    const hipError_t se = hipStreamSynchronize(f->stream);
    if (e == hipSuccess) e = se;
    if (e != hipSuccess) return fail(TR_ERR_DEVICE, std::string("tr_film_read: ") + hipGetErrorString(e));
    return TR_OK;
}

int tr_film_read(tr_film f, float* out, float* steady)
{
    if (!f) return fail(TR_ERR_INVALID, "tr_film_read: null film");
    if (!out && !steady) return TR_OK;
    const size_t W = size_t(f->desc.width), T = size_t(f->desc.time_bins), P = size_t(f->rows) * W;
    if (!f->layers) return tr_film_read_rows(f, 0, f->rows, out, steady);
    std::vector<float> s(f->n_elems), c(f->n_elems);
    HIP_TRY(hipMemcpyAsync(s.data(), f->d_sum, f->n_elems * sizeof(float), hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipMemcpyAsync(c.data(), f->d_comp, f->n_elems * sizeof(float), hipMemcpyDeviceToHost, f->stream));
    HIP_TRY(hipStreamSynchronize(f->stream));
    {
        // multibounce films: each bounce film divided by its own weights; the steady image holds
        // every sample
        std::vector<float> lw(size_t(f->layers) * P), st(P * 6);
        HIP_TRY(hipMemcpyAsync(lw.data(), f->d_layer_w, lw.size() * sizeof(float), hipMemcpyDeviceToHost, f->stream));
        HIP_TRY(hipMemcpyAsync(st.data(), f->d_steady, st.size() * sizeof(float), hipMemcpyDeviceToHost, f->stream));
        HIP_TRY(hipStreamSynchronize(f->stream));
        if (out)
            for (size_t l = 0; l < size_t(f->layers); ++l)
                for (size_t p = 0; p < P; ++p) {
                    double w = lw[l * P + p];
                    for (size_t t = 0; t < T * 3; ++t) {
                        size_t i = ((l * P + p) * T) * 3 + t;
                        double v = (double(s[i]) + double(c[i])) / w;
                        out[i] = std::isfinite(v) ? float(v) : 0.0f;
                    }
                }
        if (steady)
            for (size_t p = 0; p < P; ++p)
                for (int k = 0; k < 3; ++k) {
                    double v = (double(st[p * 3 + size_t(k)]) + double(st[(P + p) * 3 + size_t(k)])) / f->weight[p];
                    steady[p * 3 + size_t(k)] = std::isfinite(v) ? float(v) : 0.0f;
                }
        return TR_OK;
    }
}

int tr_film_write(tr_film f, const char* name, const char* ext)
{
    if (!f || !name || !*name) return fail(TR_ERR_INVALID, "tr_film_write: null film or name");
    const std::string e = ext ? ext : "hdr";
    if (e != "hdr" && e != "raw")
        return fail(TR_ERR_INVALID, "Error: unsupported output file format: \"" + e + "\"");
    const int W = f->desc.width, T = f->desc.time_bins, H = f->desc.height;
    std::vector<float> img(f->n_elems), steady(size_t(f->rows) * size_t(W) * 3);
    int rc = tr_film_read(f, img.data(), steady.data());
    if (rc) return rc;
    const size_t slice = size_t(W) * size_t(T) * 3;
    std::vector<float> tr;
    char path[4096];
    // multibounce films: bounce film b is "<name>_b%02d" (b + 1) (MultibounceStreakCameraFilm.h:127-135)
    for (int l = 0; l < std::max(f->layers, 1); ++l)
    for (int r = 0; r < f->rows; ++r) {
        const int y = f->desc.scanline >= 0 ? f->desc.scanline : r;
        // a slice is an image T wide and W high; below 8 columns it is stored transposed
        const float* data = img.data() + (size_t(l) * f->rows + size_t(r)) * slice;
        int width = T, height = W;
        if (T < 8) {
            tr.resize(slice);
            for (int x = 0; x < W; ++x)
                for (int t = 0; t < T; ++t)
                    for (int c = 0; c < 3; ++c) tr[(size_t(t) * W + x) * 3 + c] = data[(size_t(x) * T + t) * 3 + c];
            data = tr.data();
            width = W, height = T;
        }
        if (f->layers) std::snprintf(path, sizeof path, "%s_b%02d_%04d.%s", name, l + 1, y, e.c_str());
        else std::snprintf(path, sizeof path, "%s_%04d.%s", name, y, e.c_str());
        rc = e == "hdr" ? tri::write_rgbe(path, data, width, height) : tri::write_raw(path, data, slice);
        if (rc) return rc;
    }
    // steady image: W x H, time-integrated radiance (only the streak row in scanline mode)
    std::vector<float> full(size_t(W) * size_t(H) * 3, 0.0f);
    for (int r = 0; r < f->rows; ++r) {
        const int y = f->desc.scanline >= 0 ? f->desc.scanline : r;
        std::copy(steady.begin() + size_t(r) * W * 3, steady.begin() + size_t(r + 1) * W * 3,
                  full.begin() + size_t(y) * W * 3);
    }
    std::snprintf(path, sizeof path, "%s.hdr", name);
    return tri::write_rgbe(path, full.data(), W, H);
}

int tr_film_device_buffers(tr_film f, float** d_sum, float** d_comp)
{
    if (!f) return fail(TR_ERR_INVALID, "tr_film_device_buffers: null film");
    if (d_sum) *d_sum = f->d_sum;
    if (d_comp) *d_comp = f->d_comp;
    return TR_OK;
}

// StreakCameraFilm::add_samples for caller-made samples: binned on the host as the reference bins
// (pos_sensor = (time - offset) / exposure, kept if 0 <= pos_sensor < T; StreakCameraFilm.h:165-171),
// then deposited by the renders' own film kernel from shadow-record-shaped staging buffers.
int tr_film_add_samples(tr_film f, const tr_film_sample* samples, uint32_t n)
{
    if (!f) return fail(TR_ERR_INVALID, "tr_film_add_samples: null film");
    if (f->layers) return fail(TR_ERR_INVALID, "tr_film_add_samples: multibounce films take samples from renders only");
    if (n == 0) return TR_OK;
    if (!samples) return fail(TR_ERR_INVALID, "tr_film_add_samples: null samples");
    if (n > (1u << 31)) return fail(TR_ERR_INVALID, "tr_film_add_samples: more than 2^31 samples in one call");
    const tr_film_desc& fd = f->desc;
    const uint32_t W = uint32_t(fd.width), T = uint32_t(fd.time_bins);
    const int row0 = fd.scanline >= 0 ? fd.scanline : 0;
    std::vector<float4> hc(n), hd(n);
    std::vector<uint8_t> hv(n);
    std::vector<double> wadd(f->weight.size(), 0.0);
    auto bits = [](uint32_t u) {
        float v;
        std::memcpy(&v, &u, sizeof v);
        return v;
    };
    for (uint32_t i = 0; i < n; ++i) {
        const tr_film_sample& q = samples[i];
        const int row = q.y - row0;
        if (q.x < 0 || uint32_t(q.x) >= W || row < 0 || row >= f->rows)
            return fail(TR_ERR_INVALID, "tr_film_add_samples: sample " + std::to_string(i) + " lies outside the film");
        const uint32_t pix = uint32_t(row) * W + uint32_t(q.x);
        wadd[pix] += double(q.weight);
        const float sensor = (q.time - fd.offset) / fd.exposure;
        const bool in_window = sensor < float(T) && sensor >= 0.0f;
        const uint32_t bin = in_window ? std::min(uint32_t(std::floor(sensor)), T - 1u) : 0u;
        hc[i] = make_float4(q.rgb[0], q.rgb[1], q.rgb[2], bits(pix));
        hd[i] = make_float4(0.0f, 0.0f, 0.0f, bits(pix * T + bin));
        hv[i] = in_window ? 1 : 0;
    }
    float4 *dc = nullptr, *dd = nullptr;
    uint8_t* dv = nullptr;
    uint32_t* dcount = nullptr;
    auto release = [&]() {
        dev_free(dc); dev_free(dd); dev_free(dv); dev_free(dcount);
    };
    int rc;
    if ((rc = dev_alloc(&dc, n)) || (rc = dev_alloc(&dd, n)) || (rc = dev_alloc(&dv, n)) || (rc = dev_alloc(&dcount, C_WORDS))) {
        release();
        return rc;
    }
    uint32_t counts[C_WORDS] = {};
    counts[shword(0)] = n;
    RenderDev R{};
    R.film = f->d_sum;
    R.film_comp = f->d_comp;
    R.W = fd.width;
    R.H = fd.height;
    R.T = fd.time_bins;
    WaveDev Wv{};
    Wv.sh_c = dc;
    Wv.sh_d = dd;
    Wv.sh_vis = dv;
    Wv.counters = dcount;
    const uint32_t pix_slots = std::min<uint32_t>(kFilmMaxSlots, kFilmLdsBytes / (T * 24u));
    const size_t lds = size_t(pix_slots) * size_t(T) * 24u;
    const uint32_t fg = std::max(1u, std::min((n + kFilmRange - 1) / kFilmRange, uint32_t(tri::compute_units()) * 4u));
    hipError_t e = hipMemcpyAsync(dc, hc.data(), size_t(n) * sizeof(float4), hipMemcpyHostToDevice, f->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dd, hd.data(), size_t(n) * sizeof(float4), hipMemcpyHostToDevice, f->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dv, hv.data(), size_t(n), hipMemcpyHostToDevice, f->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dcount, counts, sizeof counts, hipMemcpyHostToDevice, f->stream);
    if (e == hipSuccess) {
        film_kernel<false><<<fg, kFilmBlock, lds, f->stream>>>(R, Wv, 0, pix_slots, nullptr);
        e = hipGetLastError();
    }
    // the staging vectors above are read by the copies: wait before they go out of scope
    const hipError_t se = hipStreamSynchronize(f->stream);
    if (e == hipSuccess) e = se;
    release();
    if (e != hipSuccess) return fail(TR_ERR_DEVICE, std::string("tr_film_add_samples: ") + hipGetErrorString(e));
    for (size_t p = 0; p < wadd.size(); ++p) f->weight[p] += wadd[p];
    return TR_OK;
}

}  // extern "C"

namespace {

// A render job as the host prepares it.
struct HostJob {
    tr_scene scene;
    tr_film film;
    JobDev dev;          // scene view, film/camera set-up, first global path index
    uint64_t paths;
    uint32_t spp2;
    int max_bounces;
};

// Camera and film set-up of one job (PinholeCamera.h:56-80; bunnykiller.cpp:1156-1160).
int make_job(tr_scene s, const tr_camera* cam, tr_film f, const tr_render_desc* d, HostJob* out, const char* who)
{
    if (!s || !cam || !f || !d) return fail(TR_ERR_INVALID, std::string(who) + ": null argument");
    if (!s->committed) return fail(TR_ERR_STATE, std::string(who) + ": scene not committed");
    if (s->lights.empty()) return fail(TR_ERR_INVALID, std::string(who) + ": scene has no light");
    if (d->sqrt_spp <= 0 || d->n_incoming <= 0 || d->max_bounces < 0)
        return fail(TR_ERR_INVALID, std::string(who) + ": bad sampling parameters");
    const tr_film_desc& fd = f->desc;
    const uint64_t spp2 = uint64_t(d->sqrt_spp) * uint64_t(d->sqrt_spp);
    const uint64_t total = uint64_t(f->rows) * uint64_t(fd.width) * spp2 * uint64_t(d->n_incoming);
    if ((uint64_t(fd.height) * fd.width * spp2 * uint64_t(d->n_incoming)) >= (1ull << 61))
        return fail(TR_ERR_INVALID, std::string(who) + ": too many paths for the RNG stream ids");
    if (total >= (1ull << 32))
        return fail(TR_ERR_INVALID, std::string(who) + ": more than 2^32 - 1 paths in one job; split into passes");
    if (uint64_t(f->rows) * uint64_t(fd.width) >= (1ull << 24))
        return fail(TR_ERR_INVALID, std::string(who) + ": film has more than 2^24 pixels");
    if (d->max_bounces > 255) return fail(TR_ERR_INVALID, std::string(who) + ": max_bounces above 255");
    auto nrm = [](float* v) {
        float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        float inv = float(1. / double(l));
        for (int k = 0; k < 3; ++k) v[k] *= inv;
    };
    auto crs = [](const float* a, const float* b, float* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    float up[3] = {cam->up[0], cam->up[1], cam->up[2]};
    nrm(up);
    float nv[3] = {cam->focus[0] - cam->position[0], cam->focus[1] - cam->position[1], cam->focus[2] - cam->position[2]};
    nrm(nv);
    float uv[3], vv[3];
    crs(nv, up, uv);
    nrm(uv);
    crs(uv, nv, vv);
    nrm(vv);
    float fov = float(M_PI / 180. * double(cam->fov_deg) / 2);
    RenderDev R{};
    R.film = f->d_sum;
    R.film_comp = f->d_comp;
    R.W = fd.width;
    R.H = fd.height;
    R.T = fd.time_bins;
    R.row0 = fd.scanline >= 0 ? fd.scanline : 0;
    R.spp = d->sqrt_spp;
    R.spp2 = int32_t(spp2);
    R.n_inc = d->n_incoming;
    R.exposure = fd.exposure;
    R.offset = fd.offset;
    R.total_time = float(fd.time_bins) * fd.exposure + fd.offset;
    // NextEventSampling::Time: m_avg_nb_vertices_path (4) / (T * exposure / time_of_flight(1) / ior)
    // (BidirectionalPathTracing.h:131, 246-252; World.h:776-786: time_of_flight(d) = d, ior = 1)
    R.time_sampling = (d->flags & TR_RENDER_TIME_SAMPLING) ? 1 : 0;
    R.ts_lambda = 4.0f / (float(fd.time_bins) * fd.exposure / 1.0f / 1.0f);
    R.inv_inc = float(1. / double(float(d->n_incoming)));
    R.unwarp = fd.camera_unwarp ? 1 : 0;
    R.max_bounces = d->max_bounces;
    R.cpos = make_float3(cam->position[0], cam->position[1], cam->position[2]);
    R.nv = make_float3(nv[0], nv[1], nv[2]);
    R.uv = make_float3(uv[0], uv[1], uv[2]);
    R.vv = make_float3(vv[0], vv[1], vv[2]);
    R.fd = cam->view_plane_dist > 0.f ? cam->view_plane_dist : float(1. / std::tan(double(fov)));
    R.xw2 = float(double(fd.width) / 2.0);
    R.yw2 = float(double(fd.height) / 2.0);
    R.ratio = float(fd.width) / float(fd.height);
    R.inv_spp = float(1. / double(float(d->sqrt_spp)));
    R.seed = d->seed;
    R.layers = f->layers;
    R.layer_stride = uint32_t(size_t(f->rows) * size_t(fd.width) * size_t(fd.time_bins));
    R.steady = f->d_steady;
    R.steady_comp = f->d_steady ? f->d_steady + size_t(f->rows) * size_t(fd.width) * 3 : nullptr;
    R.layer_w = f->d_layer_w;
    R.cs_bits = nullptr;  // set by the launch (wavefront buffers)
    out->scene = s;
    out->film = f;
    out->dev.S = s->view();
    out->dev.R = R;
    out->dev.path_begin = 0;
    out->paths = total;
    out->spp2 = uint32_t(spp2);
    out->max_bounces = d->max_bounces;
    return TR_OK;
}

// Everything the kernels of one launch need, shared by the per-wave enqueue.
struct Launch {
    hipStream_t stream;
    WaveBuffers* wb;
    std::vector<HostJob>* jobs;
    const LaunchCfg* cfg;
    tr_guide guide;
    trg::GuideView GV;
    int compute_units;
    hipEvent_t* ev;
    bool batch;
    int bounces;
    int stack_bound;
    int trav_ext, trav_sh;  // kTrav* of the closest-hit and any-hit launches
    uint32_t pix_slots;
    size_t film_lds;
    // profiling
    std::vector<std::pair<int, int>> ev_kind;
    double kernel_ms[TR_K_COUNT] = {};
    uint32_t launches[TR_K_COUNT] = {};
    bool events_on = false;
    bool extend_only = false;  // events around the extend launches only
    bool open = false;         // begin() recorded an event that end() closes
    // events of the current extend / shadow launch, stamped by the launch itself (launch_timed)
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;

    int flush_events()
    {
        if (ev_kind.empty()) return TR_OK;
        HIP_TRY(hipEventSynchronize(ev[2 * ev_kind.back().second + 1]));
        for (auto& kv : ev_kind) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ev[2 * kv.second], ev[2 * kv.second + 1]));
            kernel_ms[kv.first] += ms;
        }
        ev_kind.clear();
        return TR_OK;
    }
    int begin(int kind)
    {
        launches[kind]++;
        open = events_on && !(extend_only && kind != TR_K_EXTEND);
        if (!open) return TR_OK;
        if (ev_kind.size() == 256) {
            int r2 = flush_events();
            if (r2) return r2;
        }
        int slot = int(ev_kind.size());
        ev_kind.push_back({kind, slot});
        if (kind == TR_K_EXTEND || kind == TR_K_SHADOW) {
            ev_start = ev[2 * slot];
            ev_stop = ev[2 * slot + 1];
        } else {
            HIP_TRY(hipEventRecord(ev[2 * slot], stream));
        }
        return TR_OK;
    }
    int end()
    {
        HIP_TRY(hipGetLastError());
        if (!open) return TR_OK;
        open = false;
        if (ev_start) {
            ev_start = ev_stop = nullptr;
            return TR_OK;
        }
        HIP_TRY(hipEventRecord(ev[2 * ev_kind.back().second + 1], stream));
        return TR_OK;
    }

    // the kernels of one wavefront: camera, then per bounce extend / shade / shadow / film
    int enqueue_wave()
    {
        int rc;
        const LaunchCfg& c = *cfg;
        const HostJob& j0 = (*jobs)[0];
        const SceneDev S0 = j0.dev.S;
        const RenderDev R0 = j0.dev.R;
        const JobDev* dj = batch ? wb->d_jobs : nullptr;
        const uint32_t nj = uint32_t(jobs->size());
        WaveDev Wv = wb->view();
        Wv.cpos = R0.cpos;
        const uint32_t wave = c.wave, cu = uint32_t(std::max(compute_units, 1));
        // persistent traversal grid: trace_per_cu blocks per CU, but at least kTraceRaysPerThread rays
        // per thread of a full wavefront (a smaller scene's wavefronts run faster on fewer, longer-lived
        // blocks: 8 rays per thread is 8 % faster than 4 or 16 on the 2-M-path media workload, no
        // change on 33-M-path wavefronts) as long as every CU keeps kTraceMinPerCu blocks (small
        // renders: 128 blocks for a 262,144-path wavefront left three quarters of the SIMDs idle)
        const uint32_t by_rays = std::max((wave + kTraceBlock * kTraceRaysPerThread - 1) / (kTraceBlock * kTraceRaysPerThread),
                                          cu * kTraceMinPerCu);
        const uint32_t tg = std::max(1u, std::min({(wave + kTraceBlock - 1) / kTraceBlock, cu * c.trace_per_cu, by_rays}));
        const uint32_t fg = std::max(1u, std::min((wave + kFilmRange - 1) / kFilmRange, cu * 4u));
        const uint32_t sblock = c.mode == kShadeGuided ? kShadeBlockGuided : kShadeBlock;
        // guided shading: 2 blocks per CU (its LDS allows no more), persistent over the wavefront, so
        // each block stages the network's 38 KB of weights once per launch instead of once per
        // 256 paths
        const uint32_t scap = cu * (c.mode == kShadeGuided ? 2u : c.shade_per_cu);
        const uint32_t sg = std::max(1u, std::min((wave + sblock - 1) / sblock, scap));
        const uint32_t cg = (wave + kBlock - 1) / kBlock;
        if ((rc = begin(TR_K_CAMERA))) return rc;
        if (batch) camera_kernel<true><<<cg, kBlock, 0, stream>>>(S0, R0, dj, nj, Wv);
        else camera_kernel<false><<<cg, kBlock, 0, stream>>>(S0, R0, dj, nj, Wv);
        if ((rc = end())) return rc;
        int cur = 0;
        for (int b = 0; b < bounces; ++b) {
            if ((rc = begin(TR_K_EXTEND))) return rc;
            if (b == 0 && !batch && c.camera_packets &&
                (trav_ext == kTravWide || trav_ext == kTravWide16 || trav_ext == kTravWideKeyed)) {
                if (trav_ext == kTravWide) TR_CAMERA_DISPATCH_T(kTravWide, stack_bound, tg, kTraceBlock, stream, S0, Wv, cur);
                else if (trav_ext == kTravWide16)
                    TR_CAMERA_DISPATCH_T(kTravWide16, stack_bound, tg, kTraceBlock, stream, S0, Wv, cur);
                else TR_CAMERA_DISPATCH_T(kTravWideKeyed, stack_bound, tg, kTraceBlock, stream, S0, Wv, cur);
            } else {
                TR_TRACE_DISPATCH(trav_ext, stack_bound, batch, extend_kernel, tg, kTraceBlock, stream, S0, dj, Wv,
                                  cur | (b == 0 ? 2 : 0));
            }
            if ((rc = end())) return rc;
            if ((rc = begin(TR_K_SHADE))) return rc;
            if (batch) {
                if (c.mode == kShadeOnb)
                    shade_kernel<kShadeOnb, true><<<sg, kShadeBlock, 0, stream>>>(S0, R0, dj, Wv, cur, GV);
                else
                    shade_kernel<kShadeRef, true><<<sg, kShadeBlock, 0, stream>>>(S0, R0, dj, Wv, cur, GV);
            } else if (c.mode == kShadeGuided) {
                shade_kernel<kShadeGuided, false><<<sg, kShadeBlockGuided, 0, stream>>>(S0, R0, dj, Wv, cur, GV);
            } else if (c.mode == kShadeOnb) {
                shade_kernel<kShadeOnb, false><<<sg, kShadeBlock, 0, stream>>>(S0, R0, dj, Wv, cur, GV);
            } else {
                shade_kernel<kShadeRef, false><<<sg, kShadeBlock, 0, stream>>>(S0, R0, dj, Wv, cur, GV);
            }
            if ((rc = end())) return rc;
            if ((rc = begin(TR_K_SHADOW))) return rc;
            TR_TRACE_DISPATCH(trav_sh, stack_bound, batch, shadow_kernel, tg, kTraceBlock, stream, S0, dj, Wv, cur ^ 1);
            if ((rc = end())) return rc;
            if ((rc = begin(TR_K_FILM))) return rc;
            if (batch)
                film_batch_kernel<<<std::max(1u, std::min((wave + kBlock - 1) / kBlock, cu * 16u)), kBlock, 0, stream>>>(dj, Wv, cur ^ 1);
            else if (R0.layers > 0)
                film_mb_kernel<<<std::max(1u, std::min((wave + kBlock - 1) / kBlock, cu * 16u)), kBlock, 0, stream>>>(R0, Wv, cur ^ 1);
            else if (guide)
                film_kernel<true><<<fg, kFilmBlock, film_lds, stream>>>(R0, Wv, cur ^ 1, pix_slots, GV.contrib);
            else
                film_kernel<false><<<fg, kFilmBlock, film_lds, stream>>>(R0, Wv, cur ^ 1, pix_slots, nullptr);
            if ((rc = end())) return rc;
            cur ^= 1;
        }
        if (!batch && R0.layers > 0) {
            if ((rc = begin(TR_K_FILM))) return rc;
            const uint32_t ncs = wave / uint32_t(R0.n_inc);
            mb_weight_kernel<<<std::max(1u, std::min((ncs + kBlock - 1) / kBlock, cu * 16u)), kBlock, 0, stream>>>(R0, Wv);
            if ((rc = end())) return rc;
        }
        return TR_OK;
    }
};

static_assert(kMaxLayersTagBits == kMbLayerTagBits, "layer tag width");

// Renders every job of `jobs` in one sequence of wavefronts on `stream`; jobs share wavefronts
// when there are several (batched kernels look up each path's job).
int run_jobs(hipStream_t stream, WaveBuffers& wb, std::vector<HostJob>& jobs, const LaunchCfg& cfg, tr_guide guide,
             int compute_units, hipEvent_t* ev, bool* ev_made, tr_render_stats* stats,
             std::chrono::steady_clock::time_point t_start)
{
    if (jobs.empty()) return fail(TR_ERR_INVALID, "render: no job");
    // every return drains the stream, so no queued copy or kernel outlives the host buffers and
    // device views it was given (the normal path synchronises as well, to read the counters)
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{stream};
    const bool batch = jobs.size() > 1;
    if (batch && guide) return fail(TR_ERR_INVALID, "render: guided renders take one job");
    uint64_t total = 0;
    int bounces = 0, stack_bound = 0;
    uint32_t max_nodes4 = 0;
    for (auto& j : jobs) {
        j.dev.path_begin = total;
        total += j.paths;
        bounces = std::max(bounces, j.max_bounces);
        stack_bound = std::max(stack_bound, cfg.wide ? j.scene->stack4 : j.scene->depth);
        max_nodes4 = std::max(max_nodes4, j.scene->n_nodes4);
    }
    // 16-bit stack entries hold node indices below 65,536
    const bool small = max_nodes4 <= 65536u && cfg.stack_fmt != 1;
    const int trav_ext = !cfg.wide ? kTravBinary : !small ? kTravWide : cfg.stack_fmt == 2 ? kTravWideKeyed
                                                                      : cfg.stack_fmt == 3 ? kTravWidePP : kTravWide16;
    const int trav_sh = !cfg.wide ? kTravBinary : !small ? kTravWide : kTravWide16;
    const bool mb = jobs[0].dev.R.layers > 0;
    for (auto& j : jobs)
        if (j.dev.R.layers > 0 && (batch || guide))
            return fail(TR_ERR_INVALID, "render: multibounce films render alone and without guiding");
    LaunchCfg c = cfg;
    c.wave = uint32_t(std::min<uint64_t>(c.wave ? c.wave : kDefaultWave, std::max<uint64_t>(total, 1)));
    if (mb) {
        // wavefronts hold whole camera samples, whose layer marks are counted per wavefront; a
        // shadow record tags its camera sample within the wavefront in 27 bits (and its layer in 5),
        // so a wavefront holds at most kMbMaxCameraSamples camera samples
        const uint32_t ni = uint32_t(jobs[0].dev.R.n_inc);
        c.wave = mb_wave(c.wave, ni);
    }
    int rc = wb.ensure(c.wave);
    if (rc) return rc;
    if (mb) {
        const uint32_t ncs = c.wave / uint32_t(jobs[0].dev.R.n_inc);
        if ((rc = wb.ensure_tags()) || (rc = wb.ensure_cs(ncs))) return rc;
        HIP_TRY(hipMemsetAsync(wb.cs_bits, 0, size_t(ncs) * sizeof(uint32_t), stream));
        jobs[0].dev.R.cs_bits = wb.cs_bits;
    }
    Launch L{};
    L.stream = stream;
    L.wb = &wb;
    L.jobs = &jobs;
    L.cfg = &c;
    L.guide = guide;
    L.compute_units = compute_units;
    L.ev = ev;
    L.batch = batch;
    L.bounces = bounces;
    L.stack_bound = stack_bound;
    L.trav_ext = trav_ext;
    L.trav_sh = trav_sh;
    if (guide) {
        const HostJob& j = jobs[0];
        if (j.max_bounces < 1) return fail(TR_ERR_INVALID, "tr_render_transient_guided: max_bounces must be >= 1");
        if ((rc = wb.ensure_tags())) return rc;
        if ((rc = trg::guide_begin_render(guide, c.wave, uint32_t(j.max_bounces), j.scene->bb_lo, j.scene->bb_hi,
                                          stream, &L.GV)))
            return rc;
    }
    if (batch) {
        if ((rc = wb.ensure_batch(uint32_t(jobs.size())))) return rc;
        // the previous render on this stream ended with a synchronisation, so the pinned table is
        // free to overwrite
        JobDev* dj = wb.h_jobs;
        for (size_t k = 0; k < jobs.size(); ++k) dj[k] = jobs[k].dev;
        dj[jobs.size()] = jobs.back().dev;
        dj[jobs.size()].path_begin = total;
        HIP_TRY(hipMemcpyAsync(wb.d_jobs, dj, (jobs.size() + 1) * sizeof(JobDev), hipMemcpyHostToDevice, stream));
    }
    if (c.prof && !c.graph) {
        if (!*ev_made) {
            for (int k = 0; k < 2 * 256; ++k) HIP_TRY(hipEventCreate(&ev[k]));
            *ev_made = true;
        }
        L.events_on = true;
        L.extend_only = c.prof_extend;
    }
    // LDS histogram slots of the film kernel: as many pixels x T bins x RGB FP64 partials as fit in
    // kFilmLdsBytes
    {
        const uint32_t T = uint32_t(jobs[0].dev.R.T);
        L.pix_slots = std::min<uint32_t>(kFilmMaxSlots, kFilmLdsBytes / (T * 24u));
        L.film_lds = size_t(L.pix_slots) * size_t(T) * 24u;
    }
    HIP_TRY(hipMemsetAsync(wb.counters, 0, C_WORDS * sizeof(uint32_t), stream));
    // pass buffers are written on this stream; order it after any pending film work
    for (auto& j : jobs) HIP_TRY(hipStreamSynchronize(j.film->stream));

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto drop_graph = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    };
    uint32_t n_waves = 0;
    if (c.graph && !guide) {
        // the whole render (every wavefront's set-up, camera and bounce kernels) as one captured
        // graph, submitted with one launch; guided renders instead replay one wavefront's kernels
        // (their Adam steps between wavefronts take host-computed bias corrections)
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (uint64_t base = 0; base < total && !rc; base += c.wave) {
            const uint32_t n = uint32_t(std::min<uint64_t>(c.wave, total - base));
            ++n_waves;
            set_wave_kernel<<<1, 1, 0, stream>>>(wb.ctl, base, n);
            rc = L.enqueue_wave();
        }
        const hipError_t ce = hipStreamEndCapture(stream, &graph);
        if (rc || ce != hipSuccess) {
            drop_graph();
            return rc ? rc : fail(TR_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            drop_graph();
            return fail(TR_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
        const hipError_t le = hipGraphLaunch(exec, stream);
        if (le != hipSuccess) {
            drop_graph();
            return fail(TR_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(le));
        }
    }
    for (uint64_t base = 0; base < total && !(c.graph && !guide); base += c.wave) {
        uint32_t n = uint32_t(std::min<uint64_t>(c.wave, total - base));
        ++n_waves;
        if (guide) {
            if ((rc = L.begin(TR_K_GUIDE)) || (rc = trg::guide_wave_begin(guide, stream)) || (rc = L.end())) {
                drop_graph();
                return rc;
            }
        }
        set_wave_kernel<<<1, 1, 0, stream>>>(wb.ctl, base, n);
        if (c.graph) {
            if (!exec) {
                uint32_t launches_before[TR_K_COUNT];
                std::memcpy(launches_before, L.launches, sizeof launches_before);
                HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                rc = L.enqueue_wave();
                hipError_t ce = hipStreamEndCapture(stream, &graph);
                if (rc || ce != hipSuccess) {
                    drop_graph();
                    return rc ? rc : fail(TR_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
                }
                hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (ie != hipSuccess) {
                    drop_graph();
                    return fail(TR_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
                }
                std::memcpy(L.launches, launches_before, sizeof launches_before);
            }
            // count the kernels each replay runs
            L.launches[TR_K_CAMERA] += 1;
            for (int k : {TR_K_EXTEND, TR_K_SHADE, TR_K_SHADOW, TR_K_FILM}) L.launches[k] += uint32_t(bounces);
            if (mb) L.launches[TR_K_FILM] += 1;
            hipError_t le = hipGraphLaunch(exec, stream);
            if (le != hipSuccess) {
                drop_graph();
                return fail(TR_ERR_DEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(le));
            }
        } else if ((rc = L.enqueue_wave())) {
            return rc;
        }
        if (guide) {
            if ((rc = L.begin(TR_K_GUIDE)) || (rc = trg::guide_wave_end(guide, stream)) || (rc = L.end())) {
                drop_graph();
                return rc;
            }
        }
    }
    uint32_t counters[C_WORDS];
    hipError_t e = hipMemcpyAsync(counters, wb.counters, sizeof counters, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    drop_graph();
    if (rc) return rc;
    if (e != hipSuccess) return fail(TR_ERR_DEVICE, std::string("render: ") + hipGetErrorString(e));
    if ((rc = L.flush_events())) return rc;
    for (auto& j : jobs)
        for (double& w : j.film->weight) w += double(j.spp2);
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->paths = total;
        auto rd64 = [&](int w) { return uint64_t(counters[w]) | (uint64_t(counters[w + 1]) << 32); };
        stats->extension_rays = rd64(C_STAT_EXT);
        stats->shadow_rays = rd64(C_STAT_SHADOW);
        for (uint32_t p = 0; p < kParts; ++p) stats->contributions += rd64(int(part_word(p, P_STAT_CONTRIB)));
        stats->waves = n_waves;
        stats->bounces_max = uint32_t(bounces);
        for (int k = 0; k < TR_K_COUNT; ++k) {
            stats->kernel_ms[k] = L.kernel_ms[k];
            stats->kernel_launches[k] = L.launches[k];
        }
        stats->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    return TR_OK;
}

int render_impl(tr_scene s, const tr_camera* cam, tr_film f, const tr_render_desc* d, tr_guide guide,
                tr_render_stats* stats)
{
    auto t_start = std::chrono::steady_clock::now();
    std::vector<HostJob> jobs(1);
    int rc = make_job(s, cam, f, d, &jobs[0], guide ? "tr_render_transient_guided" : "tr_render_transient");
    if (rc) return rc;
    LaunchCfg c = s->tune;
    if (d->wave_paths) c.wave = d->wave_paths;
    c.wide = (d->flags & TR_RENDER_BINARY_BVH) == 0;
    c.graph = (d->flags & TR_RENDER_GRAPH) != 0;
    c.prof = (d->flags & (TR_RENDER_PROFILE | TR_RENDER_PROFILE_EXTEND)) != 0;
    c.prof_extend = (d->flags & TR_RENDER_PROFILE) == 0 && (d->flags & TR_RENDER_PROFILE_EXTEND) != 0;
    if (d->flags & (TR_RENDER_STACK32 | TR_RENDER_STACK_KEYED | TR_RENDER_POSTPONE_LEAVES)) c.stack_fmt = stack_fmt_of(d->flags);
    c.mode = guide ? kShadeGuided : ((d->flags & TR_RENDER_ORTHONORMAL_FRAME) ? kShadeOnb : kShadeRef);
    c.camera_packets = (d->flags & TR_RENDER_NO_CAMERA_PACKETS) == 0;
    return run_jobs(s->stream, s->wave, jobs, c, guide, s->compute_units, s->ev, &s->ev_made, stats, t_start);
}

}  // namespace

extern "C" {

int tr_render_transient(tr_scene s, const tr_camera* cam, tr_film f, const tr_render_desc* d, tr_render_stats* stats)
{
    return render_impl(s, cam, f, d, nullptr, stats);
}

int tr_render_transient_guided(tr_scene s, const tr_camera* cam, tr_film f, const tr_render_desc* d, tr_guide g,
                               tr_render_stats* stats)
{
    if (!g) return fail(TR_ERR_INVALID, "tr_render_transient_guided: null guide");
    if (d && (d->flags & TR_RENDER_TIME_SAMPLING))
        return fail(TR_ERR_INVALID, "tr_render_transient_guided: guiding does not combine with TR_RENDER_TIME_SAMPLING");
    return render_impl(s, cam, f, d, g, stats);
}

}  // extern "C"

// =============================================================================================
// batched renders and the launch-shape autotuner

struct tr_batch_s {
    struct Entry {
        tr_scene s;
        tr_camera cam;
        tr_film f;
        tr_render_desc d;
    };
    std::vector<Entry> entries;
    hipStream_t stream = nullptr;
    WaveBuffers wave;
    hipEvent_t ev[2 * 256] = {};
    bool ev_made = false;
    int compute_units = 256;
    bool tuned = false;
    LaunchCfg tune;

    ~tr_batch_s()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        wave.release();
        if (ev_made)
            for (auto& e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

LaunchCfg launch_cfg(const tr_batch_s* b, const tr_launch_desc* L)
{
    LaunchCfg c = b->tuned ? b->tune : LaunchCfg{};
    if (L) {
        if (L->wave_paths) c.wave = L->wave_paths;
        if (L->trace_per_cu) c.trace_per_cu = L->trace_per_cu;
        if (L->shade_per_cu) c.shade_per_cu = L->shade_per_cu;
        c.wide = (L->flags & TR_RENDER_BINARY_BVH) == 0;
        c.graph = (L->flags & TR_RENDER_GRAPH) != 0;
        c.prof = (L->flags & (TR_RENDER_PROFILE | TR_RENDER_PROFILE_EXTEND)) != 0;
        c.prof_extend = (L->flags & TR_RENDER_PROFILE) == 0 && (L->flags & TR_RENDER_PROFILE_EXTEND) != 0;
        c.mode = (L->flags & TR_RENDER_ORTHONORMAL_FRAME) ? kShadeOnb : kShadeRef;
        c.stack_fmt = stack_fmt_of(L->flags);
        c.per_job = (L->flags & TR_RENDER_PER_JOB) != 0;
    }
    return c;
}

// Renders the jobs of a batch: one shared sequence of wavefronts, or (per_job) one job after another
// on the same stream, with their statistics summed.
int run_batch(tr_batch_s* b, std::vector<HostJob>& jobs, const LaunchCfg& c, tr_render_stats* stats,
              std::chrono::steady_clock::time_point t0);

// sqrt_spp_cap > 0 lowers every job's sqrt_spp to at most the cap (autotuner trials)
int batch_jobs(tr_batch b, std::vector<HostJob>* jobs, int sqrt_spp_cap)
{
    jobs->resize(b->entries.size());
    for (size_t k = 0; k < b->entries.size(); ++k) {
        const auto& e = b->entries[k];
        tr_render_desc d = e.d;
        if (sqrt_spp_cap > 0) d.sqrt_spp = std::min(d.sqrt_spp, sqrt_spp_cap);
        int rc = make_job(e.s, &e.cam, e.f, &d, &(*jobs)[k], "tr_batch_render");
        if (rc) return rc;
    }
    return TR_OK;
}

}  // namespace

extern "C" {

int tr_batch_create(tr_batch* out)
{
    if (!out) return fail(TR_ERR_INVALID, "tr_batch_create: null output");
    *out = nullptr;
    int rc = check_device();
    if (rc) return rc;
    tr_batch b = new (std::nothrow) tr_batch_s();
    if (!b) return fail(TR_ERR_NOMEM, "tr_batch_create: host allocation");
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(TR_ERR_DEVICE, "tr_batch_create: hipStreamCreate failed");
    }
    b->compute_units = tri::compute_units();
    *out = b;
    return TR_OK;
}

int tr_batch_destroy(tr_batch b)
{
    delete b;
    return TR_OK;
}

int tr_batch_add(tr_batch b, tr_scene s, const tr_camera* cam, tr_film f, const tr_render_desc* d, uint32_t* job)
{
    if (!b || !s || !cam || !f || !d) return fail(TR_ERR_INVALID, "tr_batch_add: null argument");
    for (const auto& e : b->entries)
        if (e.f == f) return fail(TR_ERR_INVALID, "tr_batch_add: every job needs its own film");
    if (b->entries.size() >= 65535u) return fail(TR_ERR_INVALID, "tr_batch_add: too many jobs");
    HostJob probe;
    int rc = make_job(s, cam, f, d, &probe, "tr_batch_add");
    if (rc) return rc;
    b->entries.push_back({s, *cam, f, *d});
    if (job) *job = uint32_t(b->entries.size() - 1);
    return TR_OK;
}

int tr_batch_clear(tr_batch b)
{
    if (!b) return fail(TR_ERR_INVALID, "tr_batch_clear: null batch");
    b->entries.clear();
    return TR_OK;
}

int tr_batch_render(tr_batch b, const tr_launch_desc* launch, tr_render_stats* stats)
{
    auto t0 = std::chrono::steady_clock::now();
    if (!b) return fail(TR_ERR_INVALID, "tr_batch_render: null batch");
    if (b->entries.empty()) return fail(TR_ERR_INVALID, "tr_batch_render: no job");
    std::vector<HostJob> jobs;
    int rc = batch_jobs(b, &jobs, 0);
    if (rc) return rc;
    return run_batch(b, jobs, launch_cfg(b, launch), stats, t0);
}

// a trial replaces the best shape only when it is faster by more than run-to-run noise
constexpr double kTuneMargin = 0.98;

int tr_autotune(tr_batch b, tr_launch_desc* best, double* ms_best)
{
    if (!b) return fail(TR_ERR_INVALID, "tr_autotune: null batch");
    if (b->entries.empty()) return fail(TR_ERR_INVALID, "tr_autotune: no job");
    std::vector<HostJob> jobs;
    int rc = batch_jobs(b, &jobs, 0);
    if (rc) return rc;
    uint64_t total = 0;
    int max_spp = 1;
    for (auto& j : jobs) total += j.paths;
    for (auto& e : b->entries) max_spp = std::max(max_spp, e.d.sqrt_spp);
    // the trials render the jobs with sqrt_spp capped so that a trial holds about as many paths as
    // the largest candidate wavefront (or the whole batch, when it is smaller): the wavefront size
    // that wins depends on how many paths are in flight
    const uint64_t kMaxWave = 1ull << 25;
    const uint64_t target = std::min<uint64_t>(total, kMaxWave);
    uint64_t trial_total = 0;
    for (int cap = 1; cap <= max_spp; ++cap) {
        if ((rc = batch_jobs(b, &jobs, cap))) return rc;
        trial_total = 0;
        for (auto& j : jobs) trial_total += j.paths;
        if (trial_total >= target) break;
    }
    LaunchCfg base = b->tuned ? b->tune : LaunchCfg{};
    // one trial: the capped jobs, best of two runs
    auto trial = [&](const LaunchCfg& c, double* ms) -> int {
        double m = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            tr_render_stats st;
            auto t0 = std::chrono::steady_clock::now();
            int r = run_batch(b, jobs, c, &st, t0);
            if (r) return r;
            m = std::min(m, st.render_ms);
        }
        *ms = m;
        return TR_OK;
    };
    // the current shape is the incumbent; stage 1: wavefront size and graph replay; stage 2: persistent
    // grid shapes; stage 3: stack formats
    LaunchCfg bestc = base;
    bestc.prof = false;
    double best_ms;
    if ((rc = trial(bestc, &best_ms))) return rc;
    std::vector<uint32_t> waves;
    for (uint32_t w : {1u << 18, 1u << 20, 1u << 22, 1u << 24, 1u << 25})
        if (w == (1u << 18) || w / 2 < trial_total) waves.push_back(w);
    for (uint32_t w : waves)
        for (bool g : {false, true}) {
            if (w == base.wave && g == base.graph) continue;
            LaunchCfg c = base;
            c.wave = w;
            c.graph = g;
            c.prof = false;
            double ms;
            if ((rc = trial(c, &ms))) return rc;
            if (ms < best_ms * kTuneMargin) best_ms = ms, bestc = c;
        }
    for (uint32_t tpc : {8u, 16u, 24u, 32u})
        for (uint32_t spc : {1u, 2u, 4u}) {
            if (tpc == bestc.trace_per_cu && spc == bestc.shade_per_cu) continue;
            LaunchCfg c = bestc;
            c.trace_per_cu = tpc;
            c.shade_per_cu = spc;
            double ms;
            if ((rc = trial(c, &ms))) return rc;
            if (ms < best_ms * kTuneMargin) best_ms = ms, bestc = c;
        }
    // stage 3: one job after another with the best shape (jobs large enough to fill the GPU on
    // their own gain the single-job kernels and the LDS-privatised film). Per-job launch overheads
    // shrink with the job sizes, so this is decided on the full jobs, one run each.
    if (jobs.size() > 1) {
        std::vector<HostJob> full;
        if ((rc = batch_jobs(b, &full, 0))) return rc;
        double ms[2];
        for (int pj = 0; pj < 2; ++pj) {
            LaunchCfg c = bestc;
            c.per_job = pj != 0;
            tr_render_stats st;
            if ((rc = run_batch(b, full, c, &st, std::chrono::steady_clock::now()))) return rc;
            ms[pj] = st.render_ms;
        }
        if (ms[1] < ms[0] * kTuneMargin) bestc.per_job = true;
    }
    // stage 4: traversal stack formats
    for (int fmt : {1, 2, 3}) {
        LaunchCfg c = bestc;
        if (!c.wide || c.stack_fmt == fmt) continue;
        c.stack_fmt = fmt;
        double ms;
        if ((rc = trial(c, &ms))) return rc;
        if (ms < best_ms * kTuneMargin) best_ms = ms, bestc = c;
    }
    // the trials accumulated into the films
    for (auto& e : b->entries)
        if ((rc = tr_film_clear(e.f))) return rc;
    b->tune = bestc;
    b->tuned = true;
    if (b->entries.size() == 1) b->entries[0].s->tune = bestc;
    if (best) {
        best->wave_paths = bestc.wave;
        best->flags = (bestc.wide ? 0u : TR_RENDER_BINARY_BVH) | (bestc.graph ? TR_RENDER_GRAPH : 0u) |
                      (bestc.mode == kShadeOnb ? TR_RENDER_ORTHONORMAL_FRAME : 0u) | stack_flags_of(bestc.stack_fmt) |
                      (bestc.per_job ? TR_RENDER_PER_JOB : 0u);
        best->trace_per_cu = bestc.trace_per_cu;
        best->shade_per_cu = bestc.shade_per_cu;
    }
    if (ms_best) *ms_best = best_ms;
    return TR_OK;
}

}  // extern "C"

namespace {

// Per-job mode: jobs of at least kPerJobMinPaths paths (enough to fill the GPU on their own) render
// one after another with the single-job kernels and LDS-privatised film; the smaller ones share one
// sequence of wavefronts, where launch overheads are paid once for all of them.
constexpr uint64_t kPerJobMinPaths = 1ull << 22;

int run_batch(tr_batch_s* b, std::vector<HostJob>& jobs, const LaunchCfg& c, tr_render_stats* stats,
              std::chrono::steady_clock::time_point t0)
{
    if (!c.per_job || jobs.size() == 1)
        return run_jobs(b->stream, b->wave, jobs, c, nullptr, b->compute_units, b->ev, &b->ev_made, stats, t0);
    std::vector<std::vector<HostJob>> groups;
    std::vector<HostJob> small;
    for (auto& j : jobs) {
        if (j.paths >= kPerJobMinPaths) groups.emplace_back(1, j);
        else small.push_back(j);
    }
    if (!small.empty()) groups.push_back(std::move(small));
    tr_render_stats total{};
    for (auto& g : groups) {
        tr_render_stats st{};
        int rc = run_jobs(b->stream, b->wave, g, c, nullptr, b->compute_units, b->ev, &b->ev_made, &st, t0);
        if (rc) return rc;
        total.paths += st.paths;
        total.extension_rays += st.extension_rays;
        total.shadow_rays += st.shadow_rays;
        total.contributions += st.contributions;
        total.waves += st.waves;
        total.bounces_max = std::max(total.bounces_max, st.bounces_max);
        for (int k = 0; k < TR_K_COUNT; ++k) {
            total.kernel_ms[k] += st.kernel_ms[k];
            total.kernel_launches[k] += st.kernel_launches[k];
        }
    }
    total.render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = total;
    return TR_OK;
}

}  // namespace
