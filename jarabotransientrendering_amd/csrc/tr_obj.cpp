// Wavefront OBJ triangle meshes, read the way the reference's World::add_triangle_mesh reads them
// (src/RayTracing/World.h:209-541):
//   - "v x y z [w]", "vt u v [w]", "vn x y z" records; "vp" and every other record are ignored;
//   - faces of exactly three corners in the forms "i", "i/t", "i//n" and "i/t/n"; OBJ indices
//     start at 1, negative indices count back from the last record read; anything else after
//     "f" (a quad, a malformed corner) is an error;
//   - the shading normal of a corner is its "vn" normalised, or the geometric normal
//     normalize(cross(p2 - p0, p1 - p0)) for corners without one (World.h:449, 490-509).
// Texture coordinates are parsed for validation only; the transient path uses none.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "tr_api.h"
#include "tr_internal.h"

namespace tri {

namespace {

struct Cursor {
    const char* p;
    void skip_ws()
    {
        while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    }
    bool at_end()
    {
        skip_ws();
        return *p == '\0';
    }
    bool number(float* out)
    {
        skip_ws();
        char* end = nullptr;
        errno = 0;
        float v = std::strtof(p, &end);
        if (end == p) return false;
        p = end;
        *out = v;
        return true;
    }
    bool integer(long* out)
    {
        char* end = nullptr;
        long v = std::strtol(p, &end, 10);
        if (end == p) return false;
        p = end;
        *out = v;
        return true;
    }
};

}  // namespace

int load_obj(const char* path, ObjMesh* out)
{
    std::ifstream in(path);
    const std::string name(path ? path : "");
    if (!in.is_open()) return fail(TR_ERR_INVALID, "Error: failed to open model file: \"" + name + "\"");
    std::vector<float> v, vn;
    size_t n_vt = 0;
    struct Corner {
        long v, t, n;
    };
    std::vector<Corner> corners;
    std::string line;
    size_t line_num = 0;
    auto bad = [&](const std::string& what) {
        return fail(TR_ERR_INVALID, "Error: reading open model file \"" + name + "\"" + what + " at line " +
                                        std::to_string(line_num));
    };
    // OBJ index -> 0-based index into a list of `count` records
    auto resolve = [&](long index, size_t count, long* out_index) {
        long k = index > 0 ? index - 1 : long(count) + index;
        if (index == 0 || k < 0 || k >= long(count)) return false;
        *out_index = k;
        return true;
    };
    while (std::getline(in, line)) {
        ++line_num;
        Cursor c{line.c_str()};
        c.skip_ws();
        if (*c.p == '#' || *c.p == '\0') continue;
        const char* key = c.p;
        size_t klen = 0;
        while (key[klen] && key[klen] != ' ' && key[klen] != '\t' && key[klen] != '\r') ++klen;
        std::string k(key, klen);
        c.p += klen;
        if (k == "v") {
            float x, y, z, w;
            if (!c.number(&x) || !c.number(&y) || !c.number(&z)) return bad("");
            c.number(&w);
            if (!c.at_end()) return bad("");
            v.insert(v.end(), {x, y, z});
        } else if (k == "vt") {
            float a, b, w;
            if (!c.number(&a) || !c.number(&b)) return bad("");
            c.number(&w);
            if (!c.at_end()) return bad("");
            ++n_vt;
        } else if (k == "vn") {
            float x, y, z;
            if (!c.number(&x) || !c.number(&y) || !c.number(&z) || !c.at_end()) return bad("");
            vn.insert(vn.end(), {x, y, z});
        } else if (k == "f") {
            // three corners of one form: 0 "i", 1 "i/t", 2 "i//n", 3 "i/t/n"; the syntax is checked
            // first, then the indices
            long raw[3][3] = {};
            int form = -1;
            for (int j = 0; j < 3; ++j) {
                c.skip_ws();
                int f = 0;
                if (!c.integer(&raw[j][0])) return bad(": invalid face");
                if (*c.p == '/') {
                    ++c.p;
                    if (*c.p == '/') {
                        ++c.p;
                        if (!c.integer(&raw[j][2])) return bad(": invalid face");
                        f = 2;
                    } else {
                        if (!c.integer(&raw[j][1])) return bad(": invalid face");
                        f = 1;
                        if (*c.p == '/') {
                            ++c.p;
                            if (!c.integer(&raw[j][2])) return bad(": invalid face");
                            f = 3;
                        }
                    }
                }
                if (form >= 0 && f != form) return bad(": invalid face");
                form = f;
            }
            if (!c.at_end()) return bad(": invalid face");  // more than three corners
            Corner tri[3];
            for (int j = 0; j < 3; ++j) {
                Corner& cr = tri[j];
                cr.n = -1;
                if (!resolve(raw[j][0], v.size() / 3, &cr.v) || ((form & 1) && !resolve(raw[j][1], n_vt, &cr.t)) ||
                    (form >= 2 && !resolve(raw[j][2], vn.size() / 3, &cr.n)))
                    return bad(": out of range index");
            }
            corners.insert(corners.end(), tri, tri + 3);
        }
        // every other record ("vp", "o", "g", "s", "usemtl", ...) is ignored
    }
    const size_t nt = corners.size() / 3;
    out->xyz = std::move(v);
    out->idx.resize(nt * 3);
    out->normals.assign(nt * 9, 0.f);
    out->has_normals = false;
    for (size_t t = 0; t < nt; ++t) {
        const float* P[3];
        for (int j = 0; j < 3; ++j) {
            out->idx[3 * t + j] = uint32_t(corners[3 * t + j].v);
            P[j] = &out->xyz[3 * size_t(corners[3 * t + j].v)];
        }
        float e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
        float e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
        float g[3] = {e2[1] * e1[2] - e2[2] * e1[1], e2[2] * e1[0] - e2[0] * e1[2], e2[0] * e1[1] - e2[1] * e1[0]};
        float gl = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        for (int j = 0; j < 3; ++j) {
            const Corner& cr = corners[3 * t + j];
            float* dst = &out->normals[9 * t + 3 * size_t(j)];
            if (cr.n >= 0) {
                const float* n = &vn[3 * size_t(cr.n)];
                float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                for (int k = 0; k < 3; ++k) dst[k] = l > 0.f ? n[k] / l : 0.f;
                out->has_normals = true;
            } else {
                for (int k = 0; k < 3; ++k) dst[k] = gl > 0.f ? g[k] / gl : 0.f;
            }
        }
    }
    return TR_OK;
}

}  // namespace tri

extern "C" int tr_obj_load(const char* path, float** xyz, uint32_t* nv, uint32_t** idx, uint32_t* nt, float** normals)
{
    if (!path || !xyz || !nv || !idx || !nt) return tri::fail(TR_ERR_INVALID, "tr_obj_load: null argument");
    tri::ObjMesh m;
    int rc = tri::load_obj(path, &m);
    if (rc) return rc;
    *nv = uint32_t(m.xyz.size() / 3);
    *nt = uint32_t(m.idx.size() / 3);
    *xyz = static_cast<float*>(std::malloc(std::max<size_t>(m.xyz.size(), 1) * sizeof(float)));
    *idx = static_cast<uint32_t*>(std::malloc(std::max<size_t>(m.idx.size(), 1) * sizeof(uint32_t)));
    if (normals) *normals = m.has_normals ? static_cast<float*>(std::malloc(m.normals.size() * sizeof(float))) : nullptr;
    if (!*xyz || !*idx || (normals && m.has_normals && !*normals)) {
        tr_obj_free(*xyz, *idx, normals ? *normals : nullptr);
        return tri::fail(TR_ERR_NOMEM, "tr_obj_load: out of host memory");
    }
    std::memcpy(*xyz, m.xyz.data(), m.xyz.size() * sizeof(float));
    std::memcpy(*idx, m.idx.data(), m.idx.size() * sizeof(uint32_t));
    if (normals && m.has_normals) std::memcpy(*normals, m.normals.data(), m.normals.size() * sizeof(float));
    return TR_OK;
}

extern "C" int tr_obj_free(float* xyz, uint32_t* idx, float* normals)
{
    std::free(xyz);
    std::free(idx);
    std::free(normals);
    return TR_OK;
}
