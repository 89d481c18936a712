// Reference command-line probe: runs the reference's own argument parser (parse_args,
// bunnykiller.cpp:980-1502) and the wiring main does before rendering (bunnykiller.cpp:1504-1600:
// default log name, camera, default light), then prints what they produced as one JSON object, so
// that fixtures pin how a command line maps to film, camera, sampler, integrator, lights, medium
// and meshes (ProgramParams.h:115-170 for the defaults).
//
// This file is TEST INFRASTRUCTURE. It is compiled only where the reference sources exist
// (oracle/build_ref.sh) and writes fixtures under tests/golden/; no product code links or calls it.
//
// How the reference's code is reached without its build system: bunnykiller.cpp is included as a
// source file with its `main` renamed. As in ref_harness.cpp, the Embree macro from bunnykiller.h is
// undefined first, which selects the reference's non-Embree scene path; there, -name-mesh builds a
// `Mesh` object (bunnykiller.cpp:1033-1041), and the probe renames that class to a subclass that
// records each mesh's file name and material before loading it with the reference's own loader.

#include "bunnykiller.h"
#undef _USE_EMBREE_

// the reference headers expect these to be included before them (bunnykiller.cpp:25-33)
#include <array>
#include <cmath>
#include <cstdio>
#include <limits>
#include <list>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Geometry/Aggregate/Mesh.h"
#include "Material/Material.h"

namespace probe {
struct MeshRecord {
    std::string file;
    const Material3D* material;
    size_t triangles;
};
std::vector<MeshRecord> meshes;
}  // namespace probe

class ProbeMesh : public Mesh {
public:
    ProbeMesh(const std::string& name, Material3D* m) : Mesh(name, m)
    {
        std::vector<Triangle> tris;
        this->push_primitives(tris);
        probe::meshes.push_back({name, m, tris.size()});
    }
};

#define Mesh ProbeMesh
#define main bunnykiller_reference_main
#include "bunnykiller.cpp"
#undef main
#undef Mesh

namespace {

// the rectangle light's frame is protected: read it through pointers to members formed in a subclass
struct RectPeek : RectangularAreaLightSource<Radiance> {
    static constexpr Vector3 RectangularAreaLightSource<Radiance>::*position = &RectPeek::m_position;
    static constexpr Vector3 RectangularAreaLightSource<Radiance>::*normal = &RectPeek::m_normal;
    static constexpr Vector3 RectangularAreaLightSource<Radiance>::*left = &RectPeek::m_left;
    static constexpr Vector3 RectangularAreaLightSource<Radiance>::*up = &RectPeek::m_up;
    static constexpr Vector2 RectangularAreaLightSource<Radiance>::*size = &RectPeek::m_size;
};

void vec(FILE* f, const char* key, double a, double b, double c, bool comma = true)
{
    std::fprintf(f, "\"%s\": [%.9g, %.9g, %.9g]%s", key, a, b, c, comma ? ", " : "");
}

// Kd of a Lambertian UberMaterial from its BRDF value: f = Kd / pi for directions above the surface
void lambertian_kd(const Material3D* m, double kd[3])
{
    Vector3 n(0., 1., 0.), wi(0., -1., 0.), wo(0., 1., 0.);
    Spectrum R;
    m->f(wi, wo, n, Vector2(0., 0.), R);
    for (int k = 0; k < 3; ++k) kd[k] = double(R[k]) * M_PI;
}

}  // namespace

int main(int argc, char* argv[])
{
    World<DIM, Radiance> w = World<DIM, Radiance>();
    UberMaterial3D m_grey;
    ProgramParams p(DIM);
    parse_args(argc, argv, p, w, m_grey);
    // what main derives before rendering (bunnykiller.cpp:1525-1530, 1603-1630, 1636-1645)
    std::string log_file = p.log_file ? p.log_file : (std::string(p.image_file) + ".log");
    if (w.nb_lights() < 1 && p.image_coordinates_vpl.empty())
        w.add_light(new PointLightSource<DIM, Radiance>(&w, p.camera_position, Radiance(1.f)));

    FILE* f = stdout;
    std::fprintf(f, "{\"film\": {\"width\": %d, \"height\": %d, \"time_bins\": %d, \"exposure\": %.9g, "
                    "\"offset\": %.9g, \"scanline\": %d, \"aspect_ratio\": %.9g, \"camera_unwarp\": %d, "
                    "\"bounce_layers\": %d, \"transient\": %d, \"image_file\": \"%s\", \"log_file\": \"%s\"}, ",
                 p.width, p.height, p.time, double(p.time_resolution), double(p.film_offset), p.scanline,
                 double(p.aspectRatioY), int(p.camera_unwarp), p.is_multibounce_streak ? p.nb_multibounces : 0,
                 int(p.transient), p.image_file, log_file.c_str());
    std::fprintf(f, "\"camera\": {");
    vec(f, "position", p.camera_position[0], p.camera_position[1], p.camera_position[2]);
    vec(f, "focus", p.camera_looking_at[0], p.camera_looking_at[1], p.camera_looking_at[2]);
    vec(f, "up", p.camera_up[0], p.camera_up[1], p.camera_up[2]);
    std::fprintf(f, "\"view_plane_dist\": %.9g}, ", double(p.camera_view_plane_dist));
    std::fprintf(f, "\"sampler\": {\"sqrt_spp\": %d}, ", p.sqrt_spp);
    std::fprintf(f, "\"integrator\": {\"n_incoming\": %u, \"max_bounces\": %d, \"time_sampling\": %d, "
                    "\"path_tracing\": %d, \"seed\": %llu, \"scattering_level_all\": %d}, ",
                 p.vol_path_tracing_samples, p.max_nb_bounces, int(p.time_sampling), int(p.path_tracing),
                 (unsigned long long)p.new_seed, int(p.scat_level == Scattering::ALL));
    std::fprintf(f, "\"lights\": [");
    for (unsigned i = 0; i < w.nb_lights(); ++i) {
        const auto* l = w.light(int(i));
        std::fprintf(f, "%s{", i ? ", " : "");
        const Radiance I = l->get_intensities();
        if (const auto* pl = dynamic_cast<const PointLightSource<DIM, Radiance>*>(l)) {
            const VectorN<DIM> q = pl->get_position();
            std::fprintf(f, "\"type\": \"point\", ");
            vec(f, "position", q[0], q[1], q[2]);
        } else if (const auto* rl = dynamic_cast<const RectangularAreaLightSource<Radiance>*>(l)) {
            std::fprintf(f, "\"type\": \"rect\", ");
            const Vector3 &q = rl->*RectPeek::position, &n = rl->*RectPeek::normal, &lf = rl->*RectPeek::left,
                          &u = rl->*RectPeek::up;
            const Vector2& sz = rl->*RectPeek::size;
            vec(f, "position", q[0], q[1], q[2]);
            vec(f, "normal", n[0], n[1], n[2]);
            vec(f, "left", lf[0], lf[1], lf[2]);
            vec(f, "up", u[0], u[1], u[2]);
            std::fprintf(f, "\"size\": [%.9g, %.9g], ", double(sz[0]), double(sz[1]));
        } else {
            std::fprintf(f, "\"type\": \"other\", ");
        }
        vec(f, "intensity", I[0], I[1], I[2], false);
        std::fprintf(f, "}");
    }
    std::fprintf(f, "], \"medium\": ");
    if (Medium<DIM>* m = w.get_medium()) {
        auto* hm = static_cast<HomogeneousMedium<DIM, PhaseFunction::Isotropic<DIM>>*>(m);
        Vector3 sa = hm->get_absorption(VectorN<DIM>(0.)).to_rgb();
        Vector3 ss = hm->get_scattering(VectorN<DIM>(0.)).to_rgb();
        std::fprintf(f, "{");
        vec(f, "sigma_a", sa[0], sa[1], sa[2]);
        vec(f, "sigma_s", ss[0], ss[1], ss[2], false);
        std::fprintf(f, "}");
    } else {
        std::fprintf(f, "null");
    }
    std::fprintf(f, ", \"meshes\": [");
    for (size_t k = 0; k < probe::meshes.size(); ++k) {
        const auto& r = probe::meshes[k];
        const Material3D* m = r.material;
        // a two-sided material evaluates its base material (TwoSided.h:52-60); the fixtures give
        // -two-sided with an explicit -lambertian
        const bool two_sided = dynamic_cast<const TwoSidedMaterial3D*>(m) != nullptr;
        double kd[3];
        const bool deflt = m == static_cast<const Material3D*>(&m_grey);
        if (deflt) {
            // the default material: a Lambertian of albedo -lambertian-rho (bunnykiller.cpp:1500-1501;
            // the reference adds it through a pointer to a local, so it is reported from p.grey)
            for (int c = 0; c < 3; ++c) kd[c] = double(p.grey[c]);
        } else {
            lambertian_kd(m, kd);
        }
        std::fprintf(f, "%s{\"file\": \"%s\", \"triangles\": %zu, \"two_sided\": %d, \"default_material\": %d, ",
                     k ? ", " : "", r.file.c_str(), r.triangles, int(two_sided), int(deflt));
        vec(f, "kd", kd[0], kd[1], kd[2], false);
        std::fprintf(f, "}");
    }
    std::fprintf(f, "]}\n");
    return 0;
}
