/*
 * CPU oracle for the transient path-tracing integrator — TEST INFRASTRUCTURE ONLY.
 * See oracle.h for the contract. Each routine cites the reference code it restates
 * (paths relative to the reference repository root, under src/).
 *
 * Pins (what checks each part against something independent of this file):
 *   pcg32                  published pcg32 known answers + the reference's own outputs
 *                          (tests/test_oracle.py)
 *   sampler, camera,       reference harness films (tests/golden/, produced by the
 *   materials, lights,     reference's own classes) in ORC_RNG_REFERENCE mode, plus
 *   medium, integrator,    closed forms (direct lighting of a plane, transmittance)
 *   film binning
 *   Moller intersector     reference harness closest-hit ids (Triangle::intersect),
 *                          brute force, analytic axis-aligned hits
 *   watertight intersector Moller ids away from edges; watertightness on shared edges
 *   BVH                    brute force over the same triangles
 */
#include "oracle.h"

#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_1_PI
#define M_1_PI 0.31830988618379067154
#endif

/* ------------------------------------------------------------------------------------------ */
/* PCG32: RandomNumbers.h:119-128 (next_int), 167-174 (seed), 83-98 (next_float)             */

void orc_pcg32_seed(orc_pcg32* r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    orc_pcg32_next(r);
    r->state += initstate;
    orc_pcg32_next(r);
}

void orc_pcg32_default(orc_pcg32* r)
{
    r->state = 0x853c49e6748fea9bULL;
    r->inc = 0xda3e39cb94b95bdbULL;
}

uint32_t orc_pcg32_next(orc_pcg32* r)
{
    uint64_t old = r->state;
    r->state = old * 6364136223846793005ULL + r->inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}

float orc_pcg32_next_float(orc_pcg32* r)
{
    union { uint32_t i; float f; } x;
    x.i = (orc_pcg32_next(r) >> 9) | 0x3f800000u;
    return x.f - 1.0f;
}

/* ------------------------------------------------------------------------------------------ */
/* small float vector helpers (LinearAlgebra/VectorN.h)                                       */

typedef struct { float x, y, z; } v3;

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 vmulv(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vneg(v3 a) { return V(-a.x, -a.y, -a.z); }
/* VectorN.h:393-409 */
static inline v3 vcross(v3 a, v3 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
/* VectorN.h:320-328 + 195-202: v *= (float)(1./length) */
static inline v3 vnorm(v3 a)
{
    float inv = (float)(1. / (double)vlen(a));
    return vmul(a, inv);
}
static inline float vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline float favg(v3 a) { return (a.x + a.y + a.z) / 3.0f; }

/* tVectorN::transform_matrix_to (VectorN.h:429-523): rotation taking `from` to `to`, applied to v */
static v3 transform_matrix_to(v3 v, v3 from, v3 to)
{
    float mtx[3][3];
    const float epsilon = 1.e-4f;
    v3 vv = vcross(from, to);
    float e = vdot(from, to);
    float f = (e < 0.0f) ? -e : e;
    if ((double)f > 1.0 - (double)epsilon) { /* T vs double literal: compared in double */
        v3 u, x;
        float c1, c2, c3;
        x.x = (from.x > 0.0f) ? from.x : -from.x;
        x.y = (from.y > 0.0f) ? from.y : -from.y;
        x.z = (from.z > 0.0f) ? from.z : -from.z;
        if (x.x < x.y) {
            if (x.x < x.z) { x.x = 1.0f; x.y = x.z = 0.0f; }
            else { x.z = 1.0f; x.x = x.y = 0.0f; }
        } else {
            if (x.y < x.z) { x.y = 1.0f; x.x = x.z = 0.0f; }
            else { x.z = 1.0f; x.x = x.y = 0.0f; }
        }
        u = vsub(x, from);
        vv = vsub(x, to);
        c1 = (float)(2.0 / (double)vdot(u, u));
        c2 = (float)(2.0 / (double)vdot(vv, vv));
        c3 = c1 * c2 * vdot(u, vv);
        float ua[3] = {u.x, u.y, u.z}, va[3] = {vv.x, vv.y, vv.z};
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++)
                mtx[i][j] = -c1 * ua[i] * ua[j] - c2 * va[i] * va[j] + c3 * va[i] * ua[j];
            mtx[i][i] += 1.0f;
        }
    } else {
        float h = (float)(1.0 / (1.0 + (double)e));
        float hvx = h * vv.x, hvz = h * vv.z;
        float hvxy = hvx * vv.y, hvxz = hvx * vv.z, hvyz = hvz * vv.y;
        mtx[0][0] = e + hvx * vv.x;
        mtx[0][1] = hvxy - vv.z;
        mtx[0][2] = hvxz + vv.y;
        mtx[1][0] = hvxy + vv.z;
        mtx[1][1] = e + h * vv.y * vv.y;
        mtx[1][2] = hvyz - vv.x;
        mtx[2][0] = hvxz - vv.y;
        mtx[2][1] = hvyz + vv.x;
        mtx[2][2] = e + hvz * vv.z;
    }
    v3 sol;
    sol.x = mtx[0][0] * v.x + mtx[0][1] * v.y + mtx[0][2] * v.z;
    sol.y = mtx[1][0] * v.x + mtx[1][1] * v.y + mtx[1][2] * v.z;
    sol.z = mtx[2][0] * v.x + mtx[2][1] * v.y + mtx[2][2] * v.z;
    return vnorm(sol);
}

/* ------------------------------------------------------------------------------------------ */
/* random streams                                                                            */

typedef struct {
    int mode;
    orc_pcg32* global;        /* ORC_RNG_REFERENCE */
    orc_pcg32 path, nee, tgt; /* ORC_RNG_PER_PATH (tgt: the time-sampling light point) */
} rngs;

static inline float rnd_path(rngs* r)
{
    return r->mode == ORC_RNG_REFERENCE ? orc_pcg32_next_float(r->global) : orc_pcg32_next_float(&r->path);
}
static inline float rnd_tgt(rngs* r)
{
    return r->mode == ORC_RNG_REFERENCE ? orc_pcg32_next_float(r->global) : orc_pcg32_next_float(&r->tgt);
}
static inline float rnd_nee(rngs* r)
{
    return r->mode == ORC_RNG_REFERENCE ? orc_pcg32_next_float(r->global) : orc_pcg32_next_float(&r->nee);
}
/* a draw the reference makes and never uses; consumed only in reference mode */
static inline void rnd_unused(rngs* r)
{
    if (r->mode == ORC_RNG_REFERENCE) (void)orc_pcg32_next_float(r->global);
}

/* Sampling.h:80-94 (spherical) */
static v3 sample_spherical(float e1, float e2, float* pdf)
{
    float phi = (float)(2. * M_PI * e1);
    float cos_theta = (float)(1. - 2. * e2);
    float sin_theta = (float)sqrt(1. - (double)(cos_theta * cos_theta));
    *pdf = (float)(.25 * M_1_PI);
    return V(cosf(phi) * sin_theta, cos_theta, sinf(phi) * sin_theta);
}

/* Sampling.h:131-143 (cosine_weighted, 3D) */
static v3 sample_cosine(float e1, float e2, float* pdf)
{
    float phi = (float)(2. * M_PI * e1);
    float cos_theta = sqrtf(e2);
    float sin_theta = (float)sqrt(1. - (double)(cos_theta * cos_theta));
    *pdf = (float)(cos_theta * M_1_PI);
    return V(cosf(phi) * sin_theta, cos_theta, sinf(phi) * sin_theta);
}

/* ------------------------------------------------------------------------------------------ */
/* scene preprocessing and intersection                                                       */

typedef struct {
    v3 v0, e1, e2, Ng; /* Embree Triangle4 layout: e1 = v0-v1, e2 = v2-v0, Ng = cross(e1,e2) (triangle.h:62-64) */
    v3 p0, p1, p2;
    v3 shading_n;      /* normalize(cross(p2-p0, p1-p0)) (World.h:449) */
} tri_t;

typedef struct {
    float lo[3], hi[3];
    int32_t left, count; /* inner: count == 0, left = first child (right = left+1); leaf: left = first prim */
} bvh_node;

typedef struct {
    const orc_scene* s;
    tri_t* tris;
    bvh_node* nodes;
    int32_t* prim;
    int32_t n_nodes;
} world_t;

static void build_tris(world_t* w)
{
    const orc_scene* s = w->s;
    w->tris = (tri_t*)malloc(sizeof(tri_t) * (size_t)(s->n_tris > 0 ? s->n_tris : 1));
    for (int i = 0; i < s->n_tris; ++i) {
        const float* t = s->tris + 9 * (size_t)i;
        tri_t* T = &w->tris[i];
        T->p0 = V(t[0], t[1], t[2]);
        T->p1 = V(t[3], t[4], t[5]);
        T->p2 = V(t[6], t[7], t[8]);
        T->v0 = T->p0;
        T->e1 = vsub(T->p0, T->p1);
        T->e2 = vsub(T->p2, T->p0);
        T->Ng = vcross(T->e1, T->e2);
        T->shading_n = vnorm(vcross(vsub(T->p2, T->p0), vsub(T->p1, T->p0)));
    }
}

/* Embree MoellerTrumboreIntersector1::intersect (triangle_intersector_moeller.h:75-110) with the
 * epilog's t = T / |den| (intersector_epilog.h:194-270; Embree uses rcp with one Newton step,
 * exact division here). Returns 1 and updates *tfar when hit in (tnear, *tfar]. */
static inline int isect_moller(const tri_t* T, v3 O, v3 D, float tnear, float* tfar)
{
    v3 C = vsub(T->v0, O);
    v3 R = vcross(D, C);
    float den = vdot(T->Ng, D);
    float absDen = fabsf(den);
    float sgn = den < 0.0f ? -1.0f : 1.0f;
    float U = vdot(R, T->e2) * sgn;
    float Vv = vdot(R, T->e1) * sgn;
    if (!(den != 0.0f && U >= 0.0f && Vv >= 0.0f && U + Vv <= absDen)) return 0;
    float Tt = vdot(T->Ng, C) * sgn;
    if (!(absDen * tnear < Tt && Tt <= absDen * (*tfar))) return 0;
    *tfar = Tt / absDen;
    return 1;
}

/* Watertight ray/triangle test: Woop, Benthin, Wald, "Watertight Ray/Triangle
 * Intersection", JCGT 2(1), 2013 (Sec. 4 algorithm, with the double-precision
 * fallback for zero edge functions and backface culling off). */
static inline int isect_watertight(const tri_t* T, v3 O, v3 D, float tnear, float* tfar)
{
    float ad[3] = {fabsf(D.x), fabsf(D.y), fabsf(D.z)};
    int kz = (ad[0] > ad[1]) ? (ad[0] > ad[2] ? 0 : 2) : (ad[1] > ad[2] ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (vget(D, kz) < 0.0f) { int tmp = kx; kx = ky; ky = tmp; }
    float Sz = 1.0f / vget(D, kz);
    float Sx = vget(D, kx) * Sz;
    float Sy = vget(D, ky) * Sz;
    v3 A = vsub(T->p0, O), B = vsub(T->p1, O), C = vsub(T->p2, O);
    float Ax = vget(A, kx) - Sx * vget(A, kz), Ay = vget(A, ky) - Sy * vget(A, kz);
    float Bx = vget(B, kx) - Sx * vget(B, kz), By = vget(B, ky) - Sy * vget(B, kz);
    float Cx = vget(C, kx) - Sx * vget(C, kz), Cy = vget(C, ky) - Sy * vget(C, kz);
    float U = Cx * By - Cy * Bx;
    float Vv = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        Vv = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return 0;
    float det = U + Vv + W;
    if (det == 0.0f) return 0;
    float Az = Sz * vget(A, kz), Bz = Sz * vget(B, kz), Cz = Sz * vget(C, kz);
    float Tt = U * Az + Vv * Bz + W * Cz;
    float t = Tt / det;
    if (!(t > tnear && t <= *tfar)) return 0;
    *tfar = t;
    return 1;
}

/* --- binned SAH BVH (oracle-local; the GPU library has its own builder) --- */

typedef struct { float lo[3], hi[3]; } aabb;
static inline void box_empty(aabb* b) { for (int k = 0; k < 3; ++k) { b->lo[k] = FLT_MAX; b->hi[k] = -FLT_MAX; } }
static inline void box_grow(aabb* b, const aabb* o)
{
    for (int k = 0; k < 3; ++k) {
        if (o->lo[k] < b->lo[k]) b->lo[k] = o->lo[k];
        if (o->hi[k] > b->hi[k]) b->hi[k] = o->hi[k];
    }
}
static inline float box_area(const aabb* b)
{
    float dx = b->hi[0] - b->lo[0], dy = b->hi[1] - b->lo[1], dz = b->hi[2] - b->lo[2];
    if (dx < 0) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

typedef struct { aabb* pb; float* pc; int32_t* prim; bvh_node* nodes; int32_t n_nodes; } builder;

static void bvh_split(builder* b, int32_t ni, int32_t first, int32_t count)
{
    bvh_node* n = &b->nodes[ni];
    aabb bb, cb;
    box_empty(&bb); box_empty(&cb);
    for (int i = first; i < first + count; ++i) {
        int p = b->prim[i];
        box_grow(&bb, &b->pb[p]);
        aabb c; for (int k = 0; k < 3; ++k) c.lo[k] = c.hi[k] = b->pc[3 * p + k];
        box_grow(&cb, &c);
    }
    memcpy(n->lo, bb.lo, sizeof bb.lo);
    memcpy(n->hi, bb.hi, sizeof bb.hi);
    if (count <= 4) { n->left = first; n->count = count; return; }
    enum { NB = 16 };
    float best = FLT_MAX; int bax = -1, bbin = 0;
    for (int ax = 0; ax < 3; ++ax) {
        float ext = cb.hi[ax] - cb.lo[ax];
        if (ext <= 0.0f) continue;
        aabb bins[NB]; int cnt[NB];
        for (int k = 0; k < NB; ++k) { box_empty(&bins[k]); cnt[k] = 0; }
        for (int i = first; i < first + count; ++i) {
            int p = b->prim[i];
            int k = (int)((b->pc[3 * p + ax] - cb.lo[ax]) / ext * NB);
            if (k >= NB) k = NB - 1;
            if (k < 0) k = 0;
            box_grow(&bins[k], &b->pb[p]); cnt[k]++;
        }
        float la[NB]; int lc[NB]; aabb acc; box_empty(&acc); int c = 0;
        for (int k = 0; k < NB; ++k) { box_grow(&acc, &bins[k]); c += cnt[k]; la[k] = box_area(&acc); lc[k] = c; }
        box_empty(&acc); c = 0;
        for (int k = NB - 1; k > 0; --k) {
            box_grow(&acc, &bins[k]); c += cnt[k];
            float cost = la[k - 1] * (float)lc[k - 1] + box_area(&acc) * (float)c;
            if (lc[k - 1] > 0 && c > 0 && cost < best) { best = cost; bax = ax; bbin = k; }
        }
    }
    int mid = first;
    if (bax >= 0) {
        float ext = cb.hi[bax] - cb.lo[bax];
        int i = first, j = first + count - 1;
        while (i <= j) {
            int p = b->prim[i];
            int k = (int)((b->pc[3 * p + bax] - cb.lo[bax]) / ext * NB);
            if (k >= NB) k = NB - 1;
            if (k < bbin) i++;
            else { int tmp = b->prim[i]; b->prim[i] = b->prim[j]; b->prim[j] = tmp; j--; }
        }
        mid = i;
    }
    if (mid == first || mid == first + count) mid = first + count / 2; /* degenerate: median by index */
    int32_t l = b->n_nodes; b->n_nodes += 2;
    b->nodes[ni].left = l; b->nodes[ni].count = 0;
    bvh_split(b, l, first, mid - first);
    bvh_split(b, l + 1, mid, first + count - mid);
}

static void build_bvh(world_t* w)
{
    int n = w->s->n_tris;
    builder b;
    b.pb = (aabb*)malloc(sizeof(aabb) * (size_t)(n > 0 ? n : 1));
    b.pc = (float*)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
    b.prim = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    b.nodes = (bvh_node*)malloc(sizeof(bvh_node) * (size_t)(2 * n + 2));
    b.n_nodes = 1;
    for (int i = 0; i < n; ++i) {
        const tri_t* T = &w->tris[i];
        for (int k = 0; k < 3; ++k) {
            float a = vget(T->p0, k), c = vget(T->p1, k), d = vget(T->p2, k);
            b.pb[i].lo[k] = fminf(a, fminf(c, d));
            b.pb[i].hi[k] = fmaxf(a, fmaxf(c, d));
            b.pc[3 * i + k] = 0.5f * (b.pb[i].lo[k] + b.pb[i].hi[k]);
        }
        b.prim[i] = i;
    }
    if (n > 0) bvh_split(&b, 0, 0, n);
    else { memset(&b.nodes[0], 0, sizeof(bvh_node)); }
    free(b.pb); free(b.pc);
    w->nodes = b.nodes; w->prim = b.prim; w->n_nodes = b.n_nodes;
}

static inline int slab(const bvh_node* n, v3 O, v3 inv, float tnear, float tfar)
{
    float t0 = tnear, t1 = tfar;
    float o[3] = {O.x, O.y, O.z}, id[3] = {inv.x, inv.y, inv.z};
    for (int k = 0; k < 3; ++k) {
        /* inv is finite (zero direction components are clamped), so no 0 * inf */
        float a = (n->lo[k] - o[k]) * id[k], b = (n->hi[k] - o[k]) * id[k];
        float lo = a < b ? a : b, hi = a < b ? b : a;
        if (lo > t0) t0 = lo;
        if (hi < t1) t1 = hi;
    }
    /* three roundings per bound: widen the far bound by 1 + 2*gamma(3) (Pharr et al., PBRT 3.9.2) */
    return t0 <= t1 * 1.00000036f;
}

/* reciprocal direction for slab tests, zero components replaced by +-1e-20 */
static inline float safe_rcp(float d)
{
    if (fabsf(d) < 1e-20f) d = copysignf(1e-20f, d);
    return 1.0f / d;
}

/* closest hit, method = ORC_ISECT_*; returns triangle index or -1; *t is updated */
static int trace(const world_t* w, v3 O, v3 D, float tnear, float* t, int method, int use_bvh, int any)
{
    int hit = -1;
    const tri_t* tris = w->tris;
    if (!use_bvh) {
        for (int i = 0; i < w->s->n_tris; ++i) {
            float tt = *t;
            int h = method == ORC_ISECT_MOLLER ? isect_moller(&tris[i], O, D, tnear, &tt)
                                               : isect_watertight(&tris[i], O, D, tnear, &tt);
            /* equal-distance ties resolve to the lower triangle index */
            if (h && (tt < *t || hit < 0)) { *t = tt; hit = i; if (any) return hit; }
        }
        return hit;
    }
    if (w->s->n_tris == 0) return -1;
    v3 inv = V(safe_rcp(D.x), safe_rcp(D.y), safe_rcp(D.z));
    int32_t stack[128]; int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const bvh_node* n = &w->nodes[stack[--sp]];
        if (!slab(n, O, inv, tnear, *t)) continue;
        if (n->count > 0) {
            for (int i = n->left; i < n->left + n->count; ++i) {
                int p = w->prim[i];
                float tt = *t;
                int h = method == ORC_ISECT_MOLLER ? isect_moller(&tris[p], O, D, tnear, &tt)
                                                   : isect_watertight(&tris[p], O, D, tnear, &tt);
                /* equal-distance ties resolve to the lower triangle index, independent of order */
                if (h && (tt < *t || hit < 0 || p < hit)) { *t = tt; hit = p; if (any) return hit; }
            }
        } else {
            stack[sp++] = n->left + 1;
            stack[sp++] = n->left;
        }
    }
    return hit;
}

int orc_intersect(const orc_scene* s, const float* rays, int32_t n, float tnear, int32_t method,
                  int32_t use_bvh, int32_t* ids, float* t)
{
    world_t w; memset(&w, 0, sizeof w); w.s = s;
    build_tris(&w);
    if (use_bvh) build_bvh(&w);
    for (int i = 0; i < n; ++i) {
        const float* r = rays + 6 * (size_t)i;
        float tf = INFINITY;
        int h = trace(&w, V(r[0], r[1], r[2]), V(r[3], r[4], r[5]), tnear, &tf, method, use_bvh, 0);
        ids[i] = h;
        t[i] = h >= 0 ? tf : -1.0f;
    }
    free(w.tris); free(w.nodes); free(w.prim);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* lights (World.h:190-199, PointLightSource.h:142-171, RectangularAreaLightSource.h:66-80,   */
/* 222-248)                                                                                   */

typedef struct { v3 pos, left, up, normal; float intensity, inv_area; } rect_t;

static rect_t rect_setup(const float* r)
{
    rect_t L;
    L.pos = V(r[0], r[1], r[2]);
    v3 tgt = V(r[3], r[4], r[5]);
    L.normal = vnorm(vsub(tgt, L.pos));   /* (dir - pos).normalized() (bunnykiller.cpp:1258) */
    float area = r[6] * r[7];
    L.inv_area = (float)(1. / (double)area);
    v3 up = (fabs((double)vdot(L.normal, V(0, 1, 0))) < 0.9999) ? V(0, 1, 0) : V(0, 0, 1);
    v3 left = vcross(L.normal, up);
    up = vcross(L.normal, left);
    L.left = vmul(left, r[6]);
    L.up = vmul(up, r[7]);
    L.intensity = r[8];
    return L;
}

typedef struct {
    world_t w;
    rect_t* rects;
    int n_lights;
    float tnear, shadow_tnear, shadow_tfar_eps;
    int has_medium;
    v3 sa, ss, st;
    float lambda;     /* mean extinction (HomogeneousMedium.h:45-58) */
    float total_time; /* T * exposure + offset (StreakCameraFilm.h:197-200) */
    int time_sampling;
    float ts_lambda;  /* NextEventSampling::Time: 4 / (T * exposure) (BidirectionalPathTracing.h:246-252) */
    float exposure;
} ctx_t;

/* World::intersects (World.h:617-639): occlusion in (tnear, dist - eps] */
static int occluded(const ctx_t* c, v3 O, v3 D, float dist)
{
    float tf = dist - c->shadow_tfar_eps;
    return trace(&c->w, O, D, c->shadow_tnear, &tf, ORC_ISECT_MOLLER, 1, 1) >= 0;
}

/* LightSource::sample(p, ...) for the picked light; returns 1 if it connects.
 * out: irradiance (RGB), dist, dir (from light to p), pdf (pl2). */
static int light_sample(ctx_t* c, rngs* R, int li, v3 p, v3* irr, float* dist, v3* dir, float* pdf)
{
    const orc_scene* s = c->w.s;
    if (li < s->n_points) {
        const float* L = s->points + 6 * (size_t)li;
        v3 lp = V(L[0], L[1], L[2]);
        v3 d = vsub(p, lp);
        float d2 = vdot(d, d);
        *dist = sqrtf(d2);
        d = vmul(d, (float)(1. / (double)*dist));
        *dir = d;
        *pdf = 1.0f;
        if (occluded(c, lp, d, *dist)) return 0;
        float att = (float)(1. / (double)d2);
        *irr = V(L[3] * att, L[4] * att, L[5] * att);
        return 1;
    }
    const rect_t* L = &c->rects[li - s->n_points];
    float u = rnd_nee(R), v = rnd_nee(R);
    *pdf = L->inv_area;
    v3 lp = vadd(vadd(L->pos, vmul(L->left, u)), vmul(L->up, v));
    v3 d = vsub(p, lp);
    *dist = vlen(d);
    d = vmul(d, (float)(1. / (double)*dist));
    *dir = d;
    float dot_light = vdot(d, L->normal);
    v3 dn = vnorm(d); /* Ray3D shadow(pos, dir) normalises again */
    if (dot_light > 0.0f && !occluded(c, lp, dn, *dist)) {
        float a = L->intensity * dot_light / (*dist * *dist);
        *irr = V(a, a, a);
        return 1;
    }
    return 0;
}

/* the reference draws a light-path origin for every incoming sample in path-tracing-only mode
 * (BidirectionalPathTracing.h:1117-1118): World::sample_light's pick (1 draw), then
 * PointLightSource::sample(ls, pdf) (PointLightSource.h:173-183: the light's position, a spherical
 * direction, 2 draws) or RectangularAreaLightSource::sample(ls, pdf) (RectangularAreaLightSource.h:
 * 250-271: a uniform point m_position + m_left*u + m_up*v, then a cosine direction, 2 draws).
 * Only the position is used, as the target of time sampling (ShortestTime); in steady-state mode
 * the draws are consumed and nothing is used. */
static v3 light_path_origin(ctx_t* c, rngs* R)
{
    v3 target = V(0, 0, 0);
    if (R->mode != ORC_RNG_REFERENCE && !c->time_sampling) return target;
    float u = rnd_tgt(R);
    int li = (int)(u * (float)c->n_lights);
    if (li >= c->n_lights) li = c->n_lights - 1;
    if (li < c->w.s->n_points) {
        const float* L = c->w.s->points + 6 * (size_t)li;
        target = V(L[0], L[1], L[2]);
        rnd_unused(R); rnd_unused(R);
    } else {
        const rect_t* L = &c->rects[li - c->w.s->n_points];
        float a = rnd_tgt(R), b = rnd_tgt(R);
        target = vadd(vadd(L->pos, vmul(L->left, a)), vmul(L->up, b));
        rnd_unused(R); rnd_unused(R);
    }
    return target;
}

/* Probability::TimeAngular (SamplingTechniques.h:170-250) with eta = 1 and c = 1 / time_of_flight(1)
 * = 1, so ec = 1: the angle theta between the new direction and the direction to the target such
 * that a free flight of r followed by the straight line to the target arrives at time t. */
typedef struct { float t_a, t_b, r, modl; } time_angular;

static time_angular ta_make(float t_a, float t_b, float r, float modl)
{
    time_angular s = {t_a, t_b, r, modl};
    const float ec = 1.0f;
    if (r > modl) s.t_a = fmaxf(s.t_a, (2.0f * r - modl) * ec);
    else s.t_a = fmaxf(s.t_a, modl * ec);
    s.t_b = fminf(s.t_b, (2.0f * r + modl) * ec);
    return s;
}
static float ta_theta(const time_angular* s, float t)
{
    const double ec = 1.0;
    float cos_t = (float)(((double)(s->modl * s->modl) + 2. * (double)t * (double)s->r / ec -
                           (double)(t * t) / (ec * ec)) / (2. * (double)s->r * (double)s->modl));
    cos_t = fmaxf(-1.0f, fminf(cos_t, 1.0f));
    return acosf(cos_t);
}
static float ta_cdf_inv(const time_angular* s, float e) { return ta_theta(s, e * (s->t_b - s->t_a) + s->t_a); }
static float ta_pdf(const time_angular* s, float theta)
{
    const float ec = 1.0f;
    double den = sqrt((double)(s->r * s->r + s->modl * s->modl) - 2.0 * (double)s->r * (double)s->modl * (double)cosf(theta));
    float dtdtheta = (float)((double)(ec * s->modl * s->r * sinf(theta)) / den);
    return dtdtheta / (s->t_b - s->t_a);
}

/* ------------------------------------------------------------------------------------------ */
/* integrator: generate_path (BidirectionalPathTracing.h:486-644) + connect_paths (976-1045)    */

#define MAXV 66
typedef struct {
    v3 pos, dir, normal, f;
    float time;
    int medium;   /* vertex in medium (scatter) */
    int mat;
    int two_sided;
} vertex_t;

/* shading normal at a surface hit: the triangle's geometric normal, or for a smooth triangle its
 * corner normals interpolated with the barycentrics of p and not renormalised (World.h:589-598,
 * Triangle.h:155-159) */
static v3 hit_normal(const world_t* w, int tri, v3 p)
{
    const float* q = w->s->vn ? w->s->vn + 9 * (size_t)tri : NULL;
    int smooth = 0;
    if (q)
        for (int k = 0; k < 9; ++k) smooth |= q[k] != 0.0f;
    if (!smooth) return w->tris[tri].shading_n;
    const tri_t* T = &w->tris[tri];
    v3 e1 = vsub(T->p1, T->p0), e2 = vsub(T->p2, T->p0), d = vsub(p, T->p0);
    float d00 = vdot(e1, e1), d01 = vdot(e1, e2), d11 = vdot(e2, e2), d20 = vdot(d, e1), d21 = vdot(d, e2);
    float inv = 1.0f / (d00 * d11 - d01 * d01);
    float u = (d11 * d20 - d01 * d21) * inv, v = (d00 * d21 - d01 * d20) * inv;
    v3 n0 = V(q[0], q[1], q[2]), n1 = V(q[3], q[4], q[5]), n2 = V(q[6], q[7], q[8]);
    return vadd(vadd(vmul(n0, 1.0f - u - v), vmul(n1, u)), vmul(n2, v));
}

typedef struct { float rgb[3]; float time; int bounce; } rsample;

static void generate_path(ctx_t* c, rngs* R, v3 O, v3 D, v3 target, int max_size, vertex_t* verts, int* nv,
                          orc_stats* st)
{
    const orc_scene* s = c->w.s;
    v3 f = V(1, 1, 1);
    float pi = 1.0f, time = 0.0f;
    int media_interaction = 0, first = 1;
    int level = 0;
    int cur_tri = -1;  /* surface hit of the previous iteration */
    v3 cur_pos = O;    /* position of the previous hit / scatter */
    v3 cur_n = V(0, 0, 0), cur_din = V(0, 0, 0);  /* its shading normal and the ray that hit it */
    int cur_two = 0;
    *nv = 0;
    for (;;) {
        v3 ut = V(0, 0, 0), us = V(0, 0, 0);
        float pd = 1.0f, sampled = INFINITY;
        const float lam = c->time_sampling ? c->ts_lambda : c->lambda;
        if (c->has_medium) {
            ut = c->st; us = c->ss;
            /* sample_next_event_distance (241-265): MeanFreePath, or Time in time-sampling mode;
             * log and exp of a float are the float overloads */
            float eps = rnd_path(R);
            sampled = -logf(eps) / lam;
            pd = expf(-(sampled * lam)) * lam;
        }
        v3 fa = V(1, 1, 1);
        float pa = 1.0f;
        if (first) {
            first = 0;
        } else if (media_interaction) {
            /* sample_outgoing_ray, medium (343-398): xi, then time-angular sampling towards the
             * target (p_ts = 1, ShortestTime) or a phase-function sample (p_ts = 0: steady-state
             * mode, or the time window left after this flight is shorter than one exposure) */
            float ph = (float)(.25 * M_1_PI);
            int phase = 1;
            v3 wo = V(0, 0, 0);
            float pdf = 0.0f;
            if (c->time_sampling) {
                v3 l = vsub(target, cur_pos);
                float mod = vlen(l);
                float ta = mod * 1.0f + time;   /* time_of_flight(mod) * ior + current delay */
                float tb = c->total_time;
                time_angular sa = ta_make(ta, tb, sampled, mod);
                phase = (tb - (ta + sampled * 1.0f)) < c->exposure;
                rnd_unused(R);                  /* xi: compared with p_ts in {0, 1} */
                if (!phase) {
                    float theta = ta_cdf_inv(&sa, rnd_path(R));
                    float phi = (float)(2. * M_PI * (double)rnd_path(R));
                    v3 dir = V((float)(cos((double)phi) * sin((double)theta)), (float)cos((double)theta), (float)(sin((double)phi) * sin((double)theta)));
                    wo = transform_matrix_to(dir, V(0, 1, 0), l);
                    /* p = phase * (1 - p_ts) + pdf(theta) / (2 pi) * p_ts, divided by p_ts */
                    pdf = (float)((double)ph * 0.0 + (double)ta_pdf(&sa, theta) / (2. * M_PI) * 1.0);
                    if (pdf < 1e-5f) phase = 1;
                }
            } else {
                rnd_unused(R);                  /* xi (unused with phase sampling) */
            }
            if (phase) {
                float e1 = rnd_path(R), e2 = rnd_path(R);
                wo = sample_spherical(e1, e2, &pdf);
                wo = transform_matrix_to(wo, V(0, 1, 0), D);
            }
            fa = vmul(c->ss, ph);      /* phase * sigma_s (PhaseFunction.h:78-84 with unit time pdf) */
            pa = pdf;
            O = cur_pos;
            D = vnorm(wo);             /* Ray() normalises */
            level++;
        } else {
            /* surface: UberMaterial::sample_outgoing_ray (UberMaterial.h:95-104): bsdf pick draw,
             * Lambertian::sample_outgoing_ray (Lambertian.h:107-122) in the (tu, n, tv) frame of
             * Intersection::set_coordinate_system (Intersection.h:167-178) */
            float ub = rnd_path(R);
            (void)ub; /* one BSDF per material: index 0 */
            float e1 = rnd_path(R), e2 = rnd_path(R);
            float pdf;
            v3 l = sample_cosine(e1, e2, &pdf);
            v3 n = cur_n;
            v3 aux = (fabsf(n.x) > fabsf(n.z)) ? V(-n.y, n.x, n.z) : V(0.0f, -n.z, n.y);
            v3 tv = vnorm(aux);
            v3 tu = vcross(n, tv);
            v3 nd = vadd(vadd(vmul(tu, l.x), vmul(n, l.y)), vmul(tv, l.z));
            nd = vnorm(nd);
            /* two-sided material hit from behind: inverted frame (TwoSided.h:80-88, Intersection.h:180-185) */
            if (cur_two && vdot(vneg(cur_din), n) < 0.0f) nd = vneg(nd);
            const float* kd = s->mats + 3 * (size_t)s->tri_mat[cur_tri];
            float inv_pi = (float)M_1_PI;
            fa = V(kd[0] * inv_pi, kd[1] * inv_pi, kd[2] * inv_pi);
            fa = vmul(fa, fabsf(vdot(n, nd)));  /* f *= dot_abs(normal, dir) (BidirectionalPathTracing.h:415-417) */
            pa = pdf;
            O = cur_pos;
            D = nd;
            level++;
        }
        /* trace (World::first_intersection, World.h:566-615) */
        float thit = INFINITY;
        st->extension_rays++;
        int tri = trace(&c->w, O, D, c->tnear, &thit, ORC_ISECT_MOLLER, 1, 0);
        if (!c->has_medium && tri < 0) return;
        /* correct_next_event_distance (275-307) */
        if (!c->has_medium) {
            sampled = thit;
            media_interaction = 0;
        } else if (tri >= 0 && thit < sampled) {
            sampled = thit;
            media_interaction = 0;
            pd = expf(-(sampled * lam));
        } else {
            media_interaction = 1;
        }
        pi *= (pd * pa);
        v3 tr = V(expf(-sampled * ut.x), expf(-sampled * ut.y), expf(-sampled * ut.z));
        f = vmul(vmulv(vmulv(f, fa), tr), (float)(1.0 / (double)pi)); /* Spectrum/Real (Spectrum.h:312-320) */
        time += sampled;
        vertex_t* vx = &verts[*nv];
        v3 albedo;
        if (media_interaction) {
            cur_pos = vadd(O, vmul(D, sampled));
            vx->medium = 1; vx->mat = -1; vx->normal = V(0, 0, 0); vx->two_sided = 0;
            albedo = V(us.x / ut.x, us.y / ut.y, us.z / ut.z);
            cur_tri = -1;
        } else {
            cur_pos = vadd(O, vmul(D, sampled));
            vx->medium = 0; vx->mat = s->tri_mat[tri];
            vx->normal = hit_normal(&c->w, tri, cur_pos);
            vx->two_sided = s->mat_flags ? (s->mat_flags[vx->mat] & 1) : 0;
            cur_n = vx->normal;
            cur_two = vx->two_sided;
            cur_din = D;
            const float* kd = s->mats + 3 * (size_t)vx->mat;
            /* 1 - get_absorption = 1 - (1 - Kd) (Lambertian.h:56-59, BidirectionalPathTracing.h:632) */
            albedo = V(1.0f - (1.0f - kd[0]), 1.0f - (1.0f - kd[1]), 1.0f - (1.0f - kd[2]));
            cur_tri = tri;
        }
        vx->pos = cur_pos; vx->dir = D; vx->f = f; vx->time = time;
        (*nv)++;
        /* sample_termination (447-481): level cap, then Russian roulette on the mean albedo, or
         * in time-sampling mode the film's time length (TotalTime, no draw) */
        pi = 1.0f;
        if (level > max_size) return;
        if (c->time_sampling) {
            if (time > c->total_time) return;
        } else {
            float a = favg(albedo);
            float eps = rnd_path(R);
            if (eps > a) return;
            pi = a;
        }
        if (*nv >= MAXV) return;
    }
}

/* connect_vertex_with_light (650-692) + connect_paths (976-1045); returns #samples */
static int connect_paths(ctx_t* c, rngs* R, const vertex_t* verts, int nv, int max_size, float inv_inc,
                         rsample* out, orc_stats* st)
{
    int ns = 0;
    for (int ie = 0; ie < nv; ++ie) {
        const vertex_t* v = &verts[ie];
        if (v->time > c->total_time) return ns;
        if (ie + 1 > max_size) return ns;
        float u = rnd_nee(R);
        int li = (int)(u * (float)c->n_lights);
        float pl1 = 1.0f / (float)c->n_lights;
        v3 irr, dir; float dist, pl2;
        st->shadow_rays++;
        if (!light_sample(c, R, li, v->pos, &irr, &dist, &dir, &pl2)) continue;
        v3 wl = vneg(dir);
        v3 fe;
        if (v->medium) {
            fe = vmul(c->ss, (float)(.25 * M_1_PI));   /* Path.h:196-203 */
        } else {
            const float* kd = c->w.s->mats + 3 * (size_t)v->mat;
            float inv_pi = (float)M_1_PI;
            fe = V(kd[0] * inv_pi, kd[1] * inv_pi, kd[2] * inv_pi);
            float d = vdot(wl, v->normal);
            /* dot_clamped, dot_abs for a two-sided material (Path.h:188-193) */
            fe = vmul(fe, v->two_sided ? fabsf(d) : (d > 0.0f ? d : 0.0f));
        }
        v3 att = V(1, 1, 1);
        if (c->has_medium)                             /* compute_attenuation (Path.h:208-226) */
            att = V(expf(-c->st.x * dist), expf(-c->st.y * dist), expf(-c->st.z * dist));
        v3 fl = vmulv(fe, vmulv(irr, att));
        float p = pl1 * pl2;
        v3 val = vmulv(v->f, fl);
        float scale = (float)(1.0 / (double)p);
        rsample* o = &out[ns++];
        o->rgb[0] = val.x * scale * inv_inc;
        o->rgb[1] = val.y * scale * inv_inc;
        o->rgb[2] = val.z * scale * inv_inc;
        o->time = dist + v->time;
        o->bounce = ie;
        st->contributions++;
    }
    return ns;
}

/* ------------------------------------------------------------------------------------------ */
/* camera (PinholeCamera.h:56-108), film mapping (Film.h:222-225), sampler (StratifiedSampler.h:47-89),
 * render loop (TransientRenderer.h:130-179), streak film deposit (StreakCameraFilm.h:139-192, 280-297) */

int orc_render(const orc_scene* s, const orc_render_desc* d, double* film_sum, double* weight,
               orc_stats* stats)
{
    if (d && d->bounce_layers > 0) return -1;
    return orc_render_ex(s, d, film_sum, weight, NULL, NULL, stats);
}

int orc_render_ex(const orc_scene* s, const orc_render_desc* d, double* film_sum, double* weight,
                  double* layer_weight, double* steady, orc_stats* stats)
{
    const int B = d ? d->bounce_layers : 0;
    if (B < 0 || B > 32 || (B > 0 && !layer_weight)) return -1;
    if (!s || !d || d->width <= 0 || d->height <= 0 || d->time_bins <= 0 || d->sqrt_spp <= 0 ||
        d->n_incoming <= 0 || (s->n_points + s->n_rects) <= 0)
        return -1;
    orc_stats st; memset(&st, 0, sizeof st);
    ctx_t c; memset(&c, 0, sizeof c);
    c.w.s = s;
    build_tris(&c.w);
    build_bvh(&c.w);
    c.n_lights = s->n_points + s->n_rects;
    c.rects = (rect_t*)malloc(sizeof(rect_t) * (size_t)(s->n_rects > 0 ? s->n_rects : 1));
    for (int i = 0; i < s->n_rects; ++i) c.rects[i] = rect_setup(s->rects + 9 * (size_t)i);
    /* _SIGMA_VISIBILITY_ (bunnykiller.h:76-78) unless overridden */
    c.tnear = d->tnear > 0.0f ? d->tnear : 1e-5f;
    c.shadow_tnear = d->shadow_tnear > 0.0f ? d->shadow_tnear : 1e-5f;
    c.shadow_tfar_eps = d->shadow_tfar_eps > 0.0f ? d->shadow_tfar_eps : 1e-5f;
    c.has_medium = s->has_medium;
    if (c.has_medium) {
        c.sa = V(s->medium[0], s->medium[1], s->medium[2]);
        c.ss = V(s->medium[3], s->medium[4], s->medium[5]);
        c.st = vadd(c.sa, c.ss);
        c.lambda = favg(c.st);
    }
    c.total_time = (float)d->time_bins * d->exposure + d->offset;
    c.time_sampling = d->time_sampling != 0;
    c.exposure = d->exposure;
    {
        /* total_time / time_of_flight(1) / ior with time_of_flight(d) = d and ior = 1 (World.h:776-786) */
        float total = (float)d->time_bins * d->exposure;
        c.ts_lambda = 4.0f / (total / 1.0f / 1.0f);
    }

    /* camera */
    const float* cm = d->camera;
    v3 cpos = V(cm[0], cm[1], cm[2]), focus = V(cm[3], cm[4], cm[5]), up = vnorm(V(cm[6], cm[7], cm[8]));
    float fov = (float)(M_PI / 180. * (double)cm[9] / 2);
    float fd = (float)(1. / tan((double)fov));
    v3 nv = vnorm(vsub(focus, cpos));
    v3 uv = vnorm(vcross(nv, up));
    v3 vv = vnorm(vcross(uv, nv));

    int W = d->width, H = d->height, T = d->time_bins, spp = d->sqrt_spp;
    float inv_spp = (float)(1. / (double)(float)spp); /* Sampler.h:45 */
    float xw2 = (float)((double)W / 2.0), yw2 = (float)((double)H / 2.0);
    float ratio = (float)W / (float)H;
    int row0, row1;
    if (d->scanline >= 0) { row0 = d->scanline; row1 = d->scanline + 1; }
    else { row0 = d->row_begin; row1 = d->row_end; }
    if (row0 < 0 || row1 > H || row0 >= row1) { free(c.w.tris); free(c.w.nodes); free(c.w.prim); free(c.rects); return -2; }

    int max_size = d->max_bounces;
    int cap = max_size + 2 < MAXV ? max_size + 2 : MAXV;
    vertex_t* verts = (vertex_t*)malloc(sizeof(vertex_t) * (MAXV + 1));
    rsample* smp = (rsample*)malloc(sizeof(rsample) * (size_t)(MAXV + 1) * (size_t)d->n_incoming);
    float inv_inc = (float)(1. / (double)(float)d->n_incoming);
    (void)cap;

    orc_pcg32 global;
    orc_pcg32_seed(&global, d->seed, 0xda3e39cb94b95bdbULL);
    rngs R; memset(&R, 0, sizeof R);
    R.mode = d->rng_mode == ORC_RNG_PER_PATH ? ORC_RNG_PER_PATH : ORC_RNG_REFERENCE;
    R.global = &global;

    struct timespec ts0, ts1;
    clock_gettime(CLOCK_MONOTONIC, &ts0);
    for (int y = row0; y < row1; ++y) {
        for (int x = 0; x < W; ++x) {
            float limit_x = nextafterf((float)(x + 1), 0.0f), limit_y = nextafterf((float)(y + 1), 0.0f);
            for (int sy = 0; sy < spp; ++sy) {
                for (int sx = 0; sx < spp; ++sx) {
                    uint64_t cs = ((uint64_t)y * (uint64_t)W + (uint64_t)x) * (uint64_t)(spp * spp) +
                                  (uint64_t)(sy * spp + sx);
                    orc_pcg32 camr;
                    float jx, jy;
                    if (R.mode == ORC_RNG_REFERENCE) {
                        if (d->rng_mode == ORC_RNG_REFERENCE_PER_CS) orc_pcg32_seed(&global, d->seed, cs);
                        jx = orc_pcg32_next_float(&global);
                        jy = orc_pcg32_next_float(&global);
                    } else {
                        orc_pcg32_seed(&camr, d->seed, (1ULL << 62) | cs);
                        jx = orc_pcg32_next_float(&camr);
                        jy = orc_pcg32_next_float(&camr);
                    }
                    float shx = ((float)sx + jx) * inv_spp, shy = ((float)sy + jy) * inv_spp;
                    float px = fminf((float)x + shx, limit_x), py = fminf((float)y + shy, limit_y);
                    st.camera_samples++;
                    weight[(size_t)(y - row0) * (size_t)W + (size_t)x] += 1.0;
                    float ic0 = (px - xw2) / (float)W, ic1 = (py - yw2) / (float)H / ratio;
                    v3 dir = vmul(nv, fd);
                    dir = vadd(dir, vmul(uv, (float)(2. * (double)ic0)));
                    dir = vsub(dir, vmul(vv, (float)(2. * (double)ic1)));
                    dir = vnorm(vnorm(dir));
                    /* World::Li's extra closest hit for the unwarp distance (World.h:759-763) */
                    float rec_dist = FLT_MAX;
                    if (d->camera_unwarp) {
                        float th = INFINITY;
                        if (trace(&c.w, cpos, dir, c.tnear, &th, ORC_ISECT_MOLLER, 1, 0) >= 0) rec_dist = th;
                    }
                    int ns = 0;
                    for (int si = 0; si < d->n_incoming; ++si) {
                        if (R.mode == ORC_RNG_PER_PATH) {
                            uint64_t pid = cs * (uint64_t)d->n_incoming + (uint64_t)si;
                            orc_pcg32_seed(&R.path, d->seed, 2 * pid);
                            orc_pcg32_seed(&R.nee, d->seed, 2 * pid + 1);
                            if (c.time_sampling) orc_pcg32_seed(&R.tgt, d->seed, (3ULL << 61) | pid);
                        }
                        st.paths++;
                        v3 target = light_path_origin(&c, &R);
                        int nvx = 0;
                        generate_path(&c, &R, cpos, dir, target, max_size, verts, &nvx, &st);
                        if (nvx == 0) continue;
                        ns += connect_paths(&c, &R, verts, nvx, max_size, inv_inc, smp + ns, &st);
                    }
                    uint32_t layers_seen = 0;
                    const size_t pix = (size_t)(y - row0) * (size_t)W + (size_t)x;
                    const size_t layer_stride = (size_t)(row1 - row0) * (size_t)W * (size_t)T * 3;
                    for (int k = 0; k < ns; ++k) {
                        float* rgb = smp[k].rgb;
                        if (!isfinite(rgb[0]) || !isfinite(rgb[1]) || !isfinite(rgb[2]))
                            rgb[0] = rgb[1] = rgb[2] = 0.0f;   /* TransientRenderer.h:167-171 */
                        int layer = 0;
                        if (B > 0) {
                            layer = smp[k].bounce < B ? smp[k].bounce : B - 1;
                            layers_seen |= 1u << layer;
                        }
                        if (steady) {
                            steady[pix * 3 + 0] += rgb[0]; steady[pix * 3 + 1] += rgb[1]; steady[pix * 3 + 2] += rgb[2];
                        }
                        float cam_time = d->camera_unwarp ? smp[k].time - rec_dist : smp[k].time;
                        float pos = (cam_time - d->offset) / d->exposure;
                        if (!(pos < (float)T && pos >= 0.0f)) continue;
                        size_t bin = (size_t)floorf(pos);
                        if (bin >= (size_t)T) continue;
                        double* dst = film_sum + (size_t)layer * layer_stride + (pix * (size_t)T + bin) * 3;
                        dst[0] += rgb[0]; dst[1] += rgb[1]; dst[2] += rgb[2];
                    }
                    for (int b = 0; b < B; ++b)
                        if (layers_seen & (1u << b)) layer_weight[(size_t)b * (size_t)(row1 - row0) * (size_t)W + pix] += 1.0;
                }
            }
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &ts1);
    st.render_seconds = (double)(ts1.tv_sec - ts0.tv_sec) + 1e-9 * (double)(ts1.tv_nsec - ts0.tv_nsec);
    free(verts); free(smp);
    free(c.w.tris); free(c.w.nodes); free(c.w.prim); free(c.rects);
    if (stats) *stats = st;
    return 0;
}
