// Internal layouts shared by the host BVH builder, the host API and the HIP kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

// ---- multibounce wavefronts ------------------------------------------------------------------
// A multibounce film's shadow record tags its camera sample within the wavefront and its layer in
// 32 bits: (camera sample) << kMbLayerTagBits | layer. A wavefront therefore holds whole camera
// samples (n_inc paths each), at most kMbMaxCameraSamples - 1 of them.
constexpr uint32_t kMbLayerTagBits = 5;  // layers <= 32
constexpr uint32_t kMbMaxCameraSamples = 1u << (32 - kMbLayerTagBits);

// wavefront size for a multibounce render asked to run `wave` paths per wavefront
inline uint32_t mb_wave(uint32_t wave, uint32_t n_inc)
{
    const uint64_t cap = uint64_t(kMbMaxCameraSamples - 1) * n_inc;
    const uint64_t w = wave < cap ? wave : cap;
    const uint64_t whole = w / n_inc * n_inc;
    return uint32_t(whole > n_inc ? whole : n_inc);
}

// ---- BVH (binary, both child boxes stored in the parent; 64 bytes per node) ----------------
//   a = (L.lo.x, L.hi.x, L.lo.y, L.hi.y)
//   b = (R.lo.x, R.hi.x, R.lo.y, R.hi.y)
//   c = (L.lo.z, L.hi.z, R.lo.z, R.hi.z)
//   d = (L.index, R.index, L.count, R.count); count > 0: leaf, index = first entry in the
//       leaf-ordered triangle array; count == 0: inner node index; count < 0: empty child
//       (only in a root over fewer than two leaves), never entered.
struct alignas(16) BvhNodeHost {
    float a[4], b[4], c[4];
    int32_t d[4];
};
static_assert(sizeof(BvhNodeHost) == 64, "node layout");

// Leaf-ordered triangle record: 3 x float4 = 48 bytes.
//   p0.xyz, w = scene triangle index (as int bits); p1.xyz, w = material index; p2.xyz, w = 0
struct alignas(16) TriHost {
    float p0[4], p1[4], p2[4];
};
static_assert(sizeof(TriHost) == 48, "tri layout");

constexpr int kBvhMaxDepth = 48;   // builder guarantee; traversal stacks hold kBvhStack entries
constexpr int kBvhStack = 64;
constexpr int kLeafMax = 2;  // 2 measured 10-15% faster traversal than 4 on the 82k-triangle scene

struct BvhBuildResult {
    std::vector<BvhNodeHost> nodes;  // root = node 0 (always an inner node)
    std::vector<TriHost> tris;       // leaf order
    int depth = 0;
    double build_ms = 0.0;
};

// tris: n*9 floats (p0, p1, p2 as given), tri_mat: n material indices.
BvhBuildResult build_bvh(const float* tris, const int32_t* tri_mat, uint32_t n);

// ---- 4-wide BVH: one 128-byte node (one cache line of the GPU) holds four child boxes ---------
//   lox, hix, loy, hiy, loz, hiz: child boxes, structure-of-arrays (6 x float4)
//   child: node index (count == 0), first leaf-ordered triangle (count > 0); count -1: empty slot
struct alignas(16) Bvh4NodeHost {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    int32_t child[4];
    int32_t count[4];
};
static_assert(sizeof(Bvh4NodeHost) == 128, "node layout");

struct Bvh4 {
    std::vector<Bvh4NodeHost> nodes;  // root = node 0
    int depth = 0;                     // levels of 4-wide nodes
    int stack_bound = 0;               // entries a traversal stack may hold (3 per level)
};

// Collapses a binary BVH into a 4-wide one. collapse_bvh4 picks the binary nodes that become wide
// nodes by minimising the SAH cost (dynamic programming); collapse_bvh4_greedy opens the inner child
// of largest surface area until a node has four children or only leaves remain.
Bvh4 collapse_bvh4(const BvhBuildResult& b);
Bvh4 collapse_bvh4_greedy(const BvhBuildResult& b);

// ---- quantized 4-wide node: 64 bytes (four 16-byte words) --------------------------------------
//   word 0: origin xyz (float), w = grid step of z (a power of two, as a float)
//   word 1: child box bytes lo_x[4], hi_x[4], lo_y[4], hi_y[4]
//   word 2: lo_z[4], hi_z[4], grid steps of x and y (floats; stored whole, the traversal reads
//           them without decoding)
//   word 3: child references: index | count << 29 (count 0: inner node, 1..4: leaf triangles
//           starting at index, 7: empty slot)
// A child box is origin + q * step per plane, q a byte; lo planes round down and hi planes up, so
// every quantized box contains its child's float box (origin + q * step is one rounding: q * step is
// exact), and the traversal's slab test stays conservative.
struct alignas(16) Bvh4QNodeHost {
    float origin[3];
    float step_z;
    uint8_t q[24];
    float step_x, step_y;
    uint32_t ref[4];
};
static_assert(sizeof(Bvh4QNodeHost) == 64, "quantized node layout");
constexpr uint32_t kQEmpty = 7u;

std::vector<Bvh4QNodeHost> quantize_bvh4(const Bvh4& b4);
// Largest |origin| + 255 * step over the axes of the nodes (rounded up): bounds the magnitude of
// every plane of the quantized BVH, which the traversal's per-ray slab margins are sized by.
float qbvh_plane_bound(const std::vector<Bvh4QNodeHost>& q4);

// ---- lights ------------------------------------------------------------------------------
enum : int32_t { kLightPoint = 0, kLightRect = 1 };
struct alignas(16) LightHost {
    float pos[4];     // point: position; rect: corner
    float left[4];    // rect: m_left * size_x
    float up[4];      // rect: m_up * size_y
    float normal[4];  // rect: normal; w = 1/area
    float I[4];       // RGB intensity (rect: scalar replicated), w = type (as float)
};
static_assert(sizeof(LightHost) == 80, "light layout");

// ---- shared host helpers (defined in tr_kernels.hip) ----------------------------------------
namespace tri {
int fail(int code, const std::string& msg);  // records the message for tr_last_error, returns code
int check_device();                          // TR_OK if the current device is a gfx950
int compute_units();                         // compute units of the current device
}  // namespace tri

// ---- image output (tr_image_io.cpp) ---------------------------------------------------------
namespace tri {
void rgbe_from_float(float r, float g, float b, uint8_t out[4]);
// RGB float image, row-major [height][width][3], as a Radiance RGBE file (run-length encoded rows)
int write_rgbe(const char* path, const float* rgb, int width, int height);
int write_raw(const char* path, const float* data, size_t count);
}  // namespace tri

// ---- OBJ meshes (tr_obj.cpp) -----------------------------------------------------------------
namespace tri {
struct ObjMesh {
    std::vector<float> xyz;      // vertices, 3 floats each
    std::vector<uint32_t> idx;   // triangles, 3 vertex indices each
    std::vector<float> normals;  // per triangle corner shading normal, 9 floats per triangle
    bool has_normals = false;    // some corner had a "vn"
};
int load_obj(const char* path, ObjMesh* out);
}  // namespace tri
