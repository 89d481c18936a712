// Host-side BVH builder: binned surface-area-heuristic split (16 bins per axis) over triangle
// centroids, leaves of at most kLeafMax triangles, depth capped at kBvhMaxDepth by switching to
// object-median splits below depth 32. Replaces the BVH that Embree builds in rtcCommit
// (World::freeze, src/RayTracing/World.h:543-550) for the GPU traversal kernels.
#include "tr_internal.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>

namespace {

struct Box {
    float lo[3], hi[3];
    void reset()
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::numeric_limits<float>::infinity();
            hi[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box& o)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], o.lo[k]);
            hi[k] = std::max(hi[k], o.hi[k]);
        }
    }
    void grow(const float* p)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const
    {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.f)) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Builder {
    std::vector<Box> pbox;
    std::vector<float> pcen;  // 3 per prim
    std::vector<uint32_t> prim;
    std::vector<BvhNodeHost> nodes;
    int max_depth = 0;

    Box bounds(uint32_t first, uint32_t count, Box* centroid) const
    {
        Box b, c;
        b.reset();
        c.reset();
        for (uint32_t i = first; i < first + count; ++i) {
            b.grow(pbox[prim[i]]);
            c.grow(&pcen[3 * prim[i]]);
        }
        *centroid = c;
        return b;
    }

    // Partition prim[first, first+count) and return the split position (first < mid < first+count).
    uint32_t split(uint32_t first, uint32_t count, const Box& cb, int depth)
    {
        constexpr int NB = 16;
        int best_axis = -1, best_bin = 0;
        float best_cost = std::numeric_limits<float>::infinity();
        if (depth < 32) {
            for (int ax = 0; ax < 3; ++ax) {
                float ext = cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0.f)) continue;
                Box bins[NB];
                uint32_t cnt[NB] = {};
                for (auto& b : bins) b.reset();
                float scale = NB / ext;
                for (uint32_t i = first; i < first + count; ++i) {
                    uint32_t p = prim[i];
                    int k = std::min(NB - 1, std::max(0, int((pcen[3 * p + ax] - cb.lo[ax]) * scale)));
                    bins[k].grow(pbox[p]);
                    cnt[k]++;
                }
                float la[NB];
                uint32_t lc[NB];
                Box acc;
                acc.reset();
                uint32_t c = 0;
                for (int k = 0; k < NB; ++k) {
                    acc.grow(bins[k]);
                    c += cnt[k];
                    la[k] = acc.area();
                    lc[k] = c;
                }
                acc.reset();
                c = 0;
                for (int k = NB - 1; k > 0; --k) {
                    acc.grow(bins[k]);
                    c += cnt[k];
                    if (lc[k - 1] == 0 || c == 0) continue;
                    float cost = la[k - 1] * float(lc[k - 1]) + acc.area() * float(c);
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = ax;
                        best_bin = k;
                    }
                }
            }
        }
        if (best_axis >= 0) {
            float ext = cb.hi[best_axis] - cb.lo[best_axis];
            float scale = NB / ext;
            auto mid = std::partition(prim.begin() + first, prim.begin() + first + count, [&](uint32_t p) {
                int k = std::min(NB - 1, std::max(0, int((pcen[3 * p + best_axis] - cb.lo[best_axis]) * scale)));
                return k < best_bin;
            });
            uint32_t m = uint32_t(mid - prim.begin());
            if (m > first && m < first + count) return m;
        }
        // object median along the widest centroid axis (also bounds the depth)
        int ax = 0;
        float ext = -1.f;
        for (int k = 0; k < 3; ++k)
            if (cb.hi[k] - cb.lo[k] > ext) {
                ext = cb.hi[k] - cb.lo[k];
                ax = k;
            }
        uint32_t m = first + count / 2;
        std::nth_element(prim.begin() + first, prim.begin() + m, prim.begin() + first + count,
                         [&](uint32_t x, uint32_t y) {
                             float cx = pcen[3 * x + ax], cy = pcen[3 * y + ax];
                             return cx < cy || (cx == cy && x < y);
                         });
        return m;
    }

    // Fill child slot `side` of node `ni` with the subtree over prim[first, first+count).
    void child(int32_t ni, int side, uint32_t first, uint32_t count, int depth)
    {
        Box cb;
        Box b = bounds(first, count, &cb);
        BvhNodeHost& n = nodes[ni];
        float* xy = side == 0 ? n.a : n.b;
        xy[0] = b.lo[0];
        xy[1] = b.hi[0];
        xy[2] = b.lo[1];
        xy[3] = b.hi[1];
        n.c[side * 2 + 0] = b.lo[2];
        n.c[side * 2 + 1] = b.hi[2];
        max_depth = std::max(max_depth, depth);
        if (count <= uint32_t(kLeafMax)) {
            nodes[ni].d[side] = int32_t(first);
            nodes[ni].d[2 + side] = int32_t(count);
            return;
        }
        int32_t ci = int32_t(nodes.size());
        nodes.emplace_back();
        nodes[ni].d[side] = ci;
        nodes[ni].d[2 + side] = 0;
        uint32_t mid = split(first, count, cb, depth);
        child(ci, 0, first, mid - first, depth + 1);
        child(ci, 1, mid, first + count - mid, depth + 1);
    }
};

void empty_child(BvhNodeHost& n, int side)
{
    const float inf = std::numeric_limits<float>::infinity();
    float* xy = side == 0 ? n.a : n.b;
    xy[0] = inf;
    xy[1] = -inf;
    xy[2] = inf;
    xy[3] = -inf;
    n.c[side * 2 + 0] = inf;
    n.c[side * 2 + 1] = -inf;
    n.d[side] = 0;
    n.d[2 + side] = -1;  // empty: the traversal skips children with a negative count
}

}  // namespace

BvhBuildResult build_bvh(const float* tris, const int32_t* tri_mat, uint32_t n)
{
    auto t0 = std::chrono::steady_clock::now();
    Builder B;
    B.pbox.resize(n);
    B.pcen.resize(3 * size_t(n));
    B.prim.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const float* t = tris + 9 * size_t(i);
        B.pbox[i].reset();
        B.pbox[i].grow(t);
        B.pbox[i].grow(t + 3);
        B.pbox[i].grow(t + 6);
        for (int k = 0; k < 3; ++k) B.pcen[3 * i + k] = 0.5f * (B.pbox[i].lo[k] + B.pbox[i].hi[k]);
        B.prim[i] = i;
    }
    B.nodes.reserve(n > 0 ? 2 * size_t(n) / 3 + 4 : 4);
    B.nodes.emplace_back();
    std::memset(&B.nodes[0], 0, sizeof(BvhNodeHost));
    if (n == 0) {
        empty_child(B.nodes[0], 0);
        empty_child(B.nodes[0], 1);
    } else if (n <= uint32_t(kLeafMax)) {
        B.child(0, 0, 0, n, 1);
        empty_child(B.nodes[0], 1);
    } else {
        Box cb;
        B.bounds(0, n, &cb);
        uint32_t mid = B.split(0, n, cb, 0);
        B.child(0, 0, 0, mid, 1);
        B.child(0, 1, mid, n - mid, 1);
    }
    BvhBuildResult r;
    r.nodes = std::move(B.nodes);
    r.depth = B.max_depth;
    r.tris.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t p = B.prim[i];
        const float* t = tris + 9 * size_t(p);
        TriHost& o = r.tris[i];
        int32_t id = int32_t(p), m = tri_mat[p];
        std::memcpy(o.p0, t, 12);
        std::memcpy(&o.p0[3], &id, 4);
        std::memcpy(o.p1, t + 3, 12);
        std::memcpy(&o.p1[3], &m, 4);
        std::memcpy(o.p2, t + 6, 12);
        o.p2[3] = 0.f;
    }
    r.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

// ---- 4-wide collapse ------------------------------------------------------------------------
namespace {

struct ChildRef {
    float lo[3], hi[3];
    int32_t index, count;  // as in BvhNodeHost::d (count 0: inner, > 0: leaf, < 0: empty)
    float area() const
    {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return (dx >= 0.f) ? 2.f * (dx * dy + dy * dz + dz * dx) : 0.f;
    }
};

ChildRef child_of(const BvhNodeHost& n, int side)
{
    ChildRef r;
    const float* xy = side == 0 ? n.a : n.b;
    r.lo[0] = xy[0];
    r.hi[0] = xy[1];
    r.lo[1] = xy[2];
    r.hi[1] = xy[3];
    r.lo[2] = n.c[side * 2 + 0];
    r.hi[2] = n.c[side * 2 + 1];
    r.index = n.d[side];
    r.count = n.d[2 + side];
    return r;
}

int32_t collapse(const std::vector<BvhNodeHost>& bin, int32_t bnode, std::vector<Bvh4NodeHost>& out, int depth,
                 int* max_depth)
{
    ChildRef refs[4];
    int n = 0;
    refs[n++] = child_of(bin[bnode], 0);
    refs[n++] = child_of(bin[bnode], 1);
    while (n < 4) {
        int best = -1;
        float best_area = -1.f;
        for (int k = 0; k < n; ++k)
            if (refs[k].count == 0 && refs[k].area() > best_area) {
                best_area = refs[k].area();
                best = k;
            }
        if (best < 0) break;
        const BvhNodeHost& open = bin[refs[best].index];
        refs[best] = child_of(open, 0);
        refs[n++] = child_of(open, 1);
    }
    int32_t me = int32_t(out.size());
    out.emplace_back();
    *max_depth = std::max(*max_depth, depth);
    for (int k = 0; k < 4; ++k) {
        const float inf = std::numeric_limits<float>::infinity();
        bool used = k < n && refs[k].count >= 0;
        out[me].lox[k] = used ? refs[k].lo[0] : inf;
        out[me].hix[k] = used ? refs[k].hi[0] : -inf;
        out[me].loy[k] = used ? refs[k].lo[1] : inf;
        out[me].hiy[k] = used ? refs[k].hi[1] : -inf;
        out[me].loz[k] = used ? refs[k].lo[2] : inf;
        out[me].hiz[k] = used ? refs[k].hi[2] : -inf;
        out[me].child[k] = used ? refs[k].index : 0;
        out[me].count[k] = used ? refs[k].count : -1;
    }
    for (int k = 0; k < n; ++k)
        if (refs[k].count == 0) {
            int32_t c = collapse(bin, refs[k].index, out, depth + 1, max_depth);
            out[me].child[k] = c;  // `out` may have grown: index, do not hold references
        }
    return me;
}

}  // namespace

// ---- SAH-optimal collapse (dynamic programming over the binary tree) ---------------------------
// Which binary nodes become 4-wide nodes is chosen to minimise the surface-area cost
//   sum over wide nodes of area * kCostNode + sum over leaves of area * count * kCostTri
// (the greedy collapse above opens the largest child instead). Ylitie, Karras, Laine, "Efficient
// Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs", HPG 2017, sec. 4.1: for a binary
// node x and j slots, D(x, j) is the least cost of x's subtree placed in j slots of its parent's
// wide node: D(x, 1) = x as a wide node of its own (area(x) * kCostNode + its best 2..4-slot
// distribution), D(x, j > 1) = the best split of the j slots between x's two children. The binary
// builder's leaves (at most kLeafMax triangles) are kept as they are.
namespace {

constexpr float kCostNode = 1.0f;  // a node step of the traversal
constexpr float kCostTri = 0.4f;   // one triangle test, relative to a node step (instruction counts)

struct SahCollapse {
    const std::vector<BvhNodeHost>& bin;
    // per binary inner node: D[j-1] for j = 1..4 and the split of D(x, j): slots of the left child
    std::vector<std::array<float, 4>> D;
    std::vector<std::array<int8_t, 4>> split;

    explicit SahCollapse(const std::vector<BvhNodeHost>& b) : bin(b), D(b.size()), split(b.size())
    {
        for (auto& d : D) d.fill(-1.0f);
    }

    // cost of child reference c given k >= 1 slots (leaves and empty slots take one)
    float slot_cost(const ChildRef& c, int k)
    {
        if (c.count < 0) return 0.0f;
        if (c.count > 0) return c.area() * kCostTri * float(c.count);
        float best = std::numeric_limits<float>::infinity();
        for (int m = 1; m <= k; ++m) best = std::min(best, dist(c.index, m));
        return best;
    }
    // best number of slots m <= k for child c (0 for leaves / empty)
    int slot_choice(const ChildRef& c, int k)
    {
        if (c.count != 0) return 1;
        int bm = 1;
        float best = std::numeric_limits<float>::infinity();
        for (int m = 1; m <= k; ++m)
            if (dist(c.index, m) < best) {
                best = dist(c.index, m);
                bm = m;
            }
        return bm;
    }
    float dist(int32_t x, int j)
    {
        float& memo = D[size_t(x)][size_t(j - 1)];
        if (memo >= 0.0f) return memo;
        const ChildRef l = child_of(bin[size_t(x)], 0), r = child_of(bin[size_t(x)], 1);
        if (j == 1) {
            // x as a wide node of its own: its box plus its children in 2..4 slots
            float best = std::numeric_limits<float>::infinity();
            for (int m = 2; m <= 4; ++m) best = std::min(best, dist(x, m));
            const float dx = bin_area(x);
            memo = dx * kCostNode + best;
            split[size_t(x)][0] = 0;
            return memo;
        }
        float best = std::numeric_limits<float>::infinity();
        int8_t bs = 1;
        for (int a = 1; a < j; ++a) {
            const float c = slot_cost(l, a) + slot_cost(r, j - a);
            if (c < best) {
                best = c;
                bs = int8_t(a);
            }
        }
        memo = best;
        split[size_t(x)][size_t(j - 1)] = bs;
        return memo;
    }
    float bin_area(int32_t x) const
    {
        ChildRef l = child_of(bin[size_t(x)], 0), r = child_of(bin[size_t(x)], 1);
        ChildRef u = l;
        for (int a = 0; a < 3; ++a) {
            if (r.count >= 0) {
                u.lo[a] = l.count >= 0 ? std::min(l.lo[a], r.lo[a]) : r.lo[a];
                u.hi[a] = l.count >= 0 ? std::max(l.hi[a], r.hi[a]) : r.hi[a];
            }
        }
        return u.area();
    }
    // the child references of x's subtree placed in j slots (x opened), appended to `out`
    void gather(int32_t x, int j, std::vector<ChildRef>& out)
    {
        const ChildRef l = child_of(bin[size_t(x)], 0), r = child_of(bin[size_t(x)], 1);
        const int a = split[size_t(x)][size_t(j - 1)];
        place(l, a, out);
        place(r, j - a, out);
    }
    void place(const ChildRef& c, int k, std::vector<ChildRef>& out)
    {
        if (c.count != 0) {
            out.push_back(c);
            return;
        }
        const int m = slot_choice(c, k);
        if (m == 1) out.push_back(c);  // its own wide node
        else gather(c.index, m, out);
    }
    int32_t emit(int32_t x, std::vector<Bvh4NodeHost>& out, int depth, int* max_depth)
    {
        int bm = 2;
        float best = std::numeric_limits<float>::infinity();
        for (int m = 2; m <= 4; ++m)
            if (dist(x, m) < best) {
                best = dist(x, m);
                bm = m;
            }
        std::vector<ChildRef> refs;
        gather(x, bm, refs);
        const int32_t me = int32_t(out.size());
        out.emplace_back();
        *max_depth = std::max(*max_depth, depth);
        const float inf = std::numeric_limits<float>::infinity();
        for (int k = 0; k < 4; ++k) {
            const bool used = k < int(refs.size()) && refs[size_t(k)].count >= 0;
            const ChildRef* c = used ? &refs[size_t(k)] : nullptr;
            out[size_t(me)].lox[k] = used ? c->lo[0] : inf;
            out[size_t(me)].hix[k] = used ? c->hi[0] : -inf;
            out[size_t(me)].loy[k] = used ? c->lo[1] : inf;
            out[size_t(me)].hiy[k] = used ? c->hi[1] : -inf;
            out[size_t(me)].loz[k] = used ? c->lo[2] : inf;
            out[size_t(me)].hiz[k] = used ? c->hi[2] : -inf;
            out[size_t(me)].child[k] = used ? c->index : 0;
            out[size_t(me)].count[k] = used ? c->count : -1;
        }
        for (int k = 0; k < int(refs.size()); ++k)
            if (refs[size_t(k)].count == 0) {
                const int32_t c = emit(refs[size_t(k)].index, out, depth + 1, max_depth);
                out[size_t(me)].child[k] = c;  // `out` may have grown: index, do not hold references
            }
        return me;
    }
};

// Entries a traversal stack can hold at most: a node step pushes its entered inner children but the
// one it descends into (or all of them when it also has entered leaves, for any-hit queries), so
// along any root-to-node path at most sum(inner children - 1, or inner children next to a leaf)
// entries wait; the largest such sum over the tree.
int stack_bound(const std::vector<Bvh4NodeHost>& nodes, int32_t n)
{
    int inner = 0, best = 0;
    bool leaf = false;
    for (int k = 0; k < 4; ++k) {
        if (nodes[size_t(n)].count[k] == 0) ++inner;
        if (nodes[size_t(n)].count[k] > 0) leaf = true;
    }
    for (int k = 0; k < 4; ++k)
        if (nodes[size_t(n)].count[k] == 0) best = std::max(best, stack_bound(nodes, nodes[size_t(n)].child[k]));
    return best + (leaf ? inner : std::max(inner - 1, 0));
}

}  // namespace

Bvh4 collapse_bvh4(const BvhBuildResult& b)
{
    Bvh4 r;
    r.nodes.reserve(b.nodes.size() / 2 + 2);
    int depth = 0;
    SahCollapse sc(b.nodes);
    sc.emit(0, r.nodes, 1, &depth);
    r.depth = depth;
    r.stack_bound = stack_bound(r.nodes, 0);
    return r;
}

// The greedy collapse (open the child of largest area until four children), kept for comparison.
Bvh4 collapse_bvh4_greedy(const BvhBuildResult& b)
{
    Bvh4 r;
    r.nodes.reserve(b.nodes.size() / 2 + 2);
    int depth = 0;
    collapse(b.nodes, 0, r.nodes, 1, &depth);
    r.depth = depth;
    r.stack_bound = stack_bound(r.nodes, 0);
    return r;
}

namespace {

// float(origin + q * 2^(e - 127)) with q * 2^(e - 127) exact, as the device evaluates it
float dequant(float origin, uint32_t q, int be)
{
    return origin + float(q) * std::ldexp(1.0f, be - 127);
}

}  // namespace

std::vector<Bvh4QNodeHost> quantize_bvh4(const Bvh4& b4)
{
    std::vector<Bvh4QNodeHost> out(b4.nodes.size());
    for (size_t ni = 0; ni < b4.nodes.size(); ++ni) {
        const Bvh4NodeHost& n = b4.nodes[ni];
        Bvh4QNodeHost& q = out[ni];
        std::memset(&q, 0, sizeof(q));
        const float* lo[3] = {n.lox, n.loy, n.loz};
        const float* hi[3] = {n.hix, n.hiy, n.hiz};
        bool used[4];
        for (int k = 0; k < 4; ++k) used[k] = n.count[k] >= 0;
        int be[3];
        for (int a = 0; a < 3; ++a) {
            float o = std::numeric_limits<float>::infinity(), h = -std::numeric_limits<float>::infinity();
            for (int k = 0; k < 4; ++k)
                if (used[k]) {
                    o = std::min(o, lo[a][k]);
                    h = std::max(h, hi[a][k]);
                }
            if (!(o <= h)) o = h = 0.f;  // no used child
            q.origin[a] = o;
            // smallest grid step 2^(e-127) whose 255th plane reaches the largest hi plane
            int e = 1;
            const double ext = double(h) - double(o);
            if (ext > 0.0) e = std::max(1, std::min(254, int(std::ceil(std::log2(ext / 255.0))) + 127 - 1));
            while (e < 254 && dequant(o, 255u, e) < h) ++e;
            be[a] = e;
        }
        q.step_x = std::ldexp(1.0f, be[0] - 127);
        q.step_y = std::ldexp(1.0f, be[1] - 127);
        q.step_z = std::ldexp(1.0f, be[2] - 127);
        // byte offsets: lo_x 0, hi_x 4, lo_y 8, hi_y 12 (word 1), lo_z 16, hi_z 20 (word 2)
        const int lo_off[3] = {0, 8, 16}, hi_off[3] = {4, 12, 20};
        for (int k = 0; k < 4; ++k) {
            if (!used[k]) {
                for (int a = 0; a < 3; ++a) {
                    q.q[lo_off[a] + k] = 255;
                    q.q[hi_off[a] + k] = 0;
                }
                q.ref[k] = kQEmpty << 29;
                continue;
            }
            for (int a = 0; a < 3; ++a) {
                const float o = q.origin[a];
                const double step = std::ldexp(1.0, be[a] - 127);
                long ql = long(std::floor((double(lo[a][k]) - double(o)) / step));
                ql = std::max(0L, std::min(255L, ql));
                while (ql > 0 && dequant(o, uint32_t(ql), be[a]) > lo[a][k]) --ql;
                long qh = long(std::ceil((double(hi[a][k]) - double(o)) / step));
                qh = std::max(0L, std::min(255L, qh));
                while (qh < 255 && dequant(o, uint32_t(qh), be[a]) < hi[a][k]) ++qh;
                q.q[lo_off[a] + k] = uint8_t(ql);
                q.q[hi_off[a] + k] = uint8_t(qh);
            }
            q.ref[k] = uint32_t(n.child[k]) | (uint32_t(n.count[k]) << 29);
        }
    }
    return out;
}

float qbvh_plane_bound(const std::vector<Bvh4QNodeHost>& q4)
{
    double c = 0.0;
    for (const Bvh4QNodeHost& q : q4) {
        const double step[3] = {q.step_x, q.step_y, q.step_z};
        for (int a = 0; a < 3; ++a) c = std::max(c, std::fabs(double(q.origin[a])) + 255.0 * step[a]);
    }
    return std::nextafter(float(c * (1.0 + 1e-6)), std::numeric_limits<float>::infinity());
}
