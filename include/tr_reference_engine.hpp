/*
 * tr_reference_engine.hpp — the reference renderer's side of the drop-in boundary: a RenderEngine
 * that renders on the GPU through include/tr_api.h, for a maintainer of the reference
 * (LeksiDor/JaraboTransientRendering) to add next to src/Renderer/TransientRenderer.h.
 *
 * It is compiled against the reference's own headers (it includes them; it is not part of this
 * library's build): oracle/check_reference_engine.sh compiles it there, with Embree's vendored
 * headers, as a compile-only check.
 *
 *   GpuTransientRenderer<Radiance, RadianceAttenuation>
 *     : RenderEngine<3, Radiance, RadianceAttenuation>            (src/Renderer/RenderEngine.h:43-68)
 *     render(name, world, integrator, camera, film, sampler)       replaces TransientRenderer::render
 *                                                                  (src/Renderer/TransientRenderer.h:77-201)
 *
 * What it reads from the reference's objects (their public interfaces only):
 *   - lights: World::nb_lights / light(i) (World.h:174-203) — PointLightSource position and
 *     intensity (PointLightSource.h:40-52, LightSource.h:66-68), RectangularAreaLightSource corner,
 *     normal and size (RectangularAreaLightSource.h:47-80, read through member pointers formed in a
 *     subclass, as the members are protected);
 *   - the medium: World::get_medium (World.h:723) as HomogeneousMedium<Isotropic>
 *     get_absorption / get_scattering (HomogeneousMedium.h:45-116), as main logs it;
 *   - materials: Material::f (Material.h) of each mesh's material at normal incidence gives the
 *     Lambertian Kd (f = Kd / pi, Lambertian.h:60-72); TwoSidedMaterial (TwoSided.h:34-96) is
 *     detected by type;
 *   - the film: Film::get_width / get_height / get_time_resolution / get_exposure_time /
 *     get_time_offset (Film.h:188-220, StreakCameraFilm.h:128-215).
 * What the caller passes because the reference keeps it private: the meshes (the Embree World keeps
 * no mesh list; main records each -name-mesh file and material as it calls add_triangle_mesh), the
 * camera and the integrator / sampler settings (ProgramParams, bunnykiller.cpp:1595-1600,
 * 1751-1818) in tr_camera and tr_render_desc form.
 *
 * The result goes back into the reference's film through its own Film::add_samples
 * (StreakCameraFilm.h:139-192): one sample of weight 1 per pixel, carrying the GPU's normalised
 * value of every non-zero time bin at the bin's centre time, pixels in scanline order (the film
 * streams its slices). The film's own averaging and write() then produce the reference's files.
 */
#ifndef TR_REFERENCE_ENGINE_HPP
#define TR_REFERENCE_ENGINE_HPP

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "Film/StreakCameraFilm.h"
#include "LightSource/PointLightSource.h"
#include "LightSource/RectangularAreaLightSource.h"
#include "Material/Material.h"
#include "Material/TwoSided.h"
#include "Media/HomogeneousMedium.h"
#include "Media/PhaseFunction/IsotropicPhaseFunction.h"
#include "Renderer/RenderEngine.h"
#include "tr_api.h"

namespace tr_reference {

inline void check(int rc, const char* what)
{
    if (rc != TR_OK) throw std::runtime_error(std::string(what) + ": " + tr_last_error());
}

// One -name-mesh of the command line: its OBJ file and the material load_material returned.
struct MeshEntry {
    std::string obj_file;
    const Material3D* material;
};

template <class Radiance>
struct RectFrame : RectangularAreaLightSource<Radiance> {
    static constexpr Vector3 RectangularAreaLightSource<Radiance>::*position = &RectFrame::m_position;
    static constexpr Vector3 RectangularAreaLightSource<Radiance>::*normal = &RectFrame::m_normal;
    static constexpr Vector2 RectangularAreaLightSource<Radiance>::*size = &RectFrame::m_size;
};

// Lambertian albedo of a mesh material: pi * f at normal incidence
inline tr_material lambertian_of(const Material3D* m)
{
    Vector3 n(0., 1., 0.), wi(0., -1., 0.), wo(0., 1., 0.);
    Spectrum R;
    m->f(wi, wo, n, Vector2(0., 0.), R);
    tr_material out;
    for (int c = 0; c < 3; ++c) out.kd[c] = float(double(R[c]) * M_PI);
    return out;
}

// The GPU scene of a reference World: the caller's meshes, the World's lights and medium.
template <class Radiance>
tr_scene make_scene(const World<3, Radiance>& w, const std::vector<MeshEntry>& meshes)
{
    tr_scene s = nullptr;
    check(tr_scene_create(&s), "tr_scene_create");
    try {
        for (const MeshEntry& m : meshes) {
            const bool two_sided = dynamic_cast<const TwoSidedMaterial3D*>(m.material) != nullptr;
            const tr_material mat = lambertian_of(m.material);
            check(tr_scene_add_obj(s, m.obj_file.c_str(), &mat, two_sided ? TR_MATERIAL_TWO_SIDED : 0u, nullptr),
                  "tr_scene_add_obj");
        }
        for (unsigned i = 0; i < w.nb_lights(); ++i) {
            const LightSource<3, Radiance>* l = w.light(int(i));
            const Radiance I = l->get_intensities();
            if (const auto* pl = dynamic_cast<const PointLightSource<3, Radiance>*>(l)) {
                const VectorN<3> q = pl->get_position();
                const float pos[3] = {float(q[0]), float(q[1]), float(q[2])};
                const float ints[3] = {float(I[0]), float(I[1]), float(I[2])};
                check(tr_scene_add_point_light(s, pos, ints), "tr_scene_add_point_light");
            } else if (const auto* rl = dynamic_cast<const RectangularAreaLightSource<Radiance>*>(l)) {
                const Vector3& q = rl->*RectFrame<Radiance>::position;
                const Vector3& n = rl->*RectFrame<Radiance>::normal;
                const Vector2& sz = rl->*RectFrame<Radiance>::size;
                // tr_scene_add_rect_light takes a target point: the normal is normalize(target - corner)
                const float corner[3] = {float(q[0]), float(q[1]), float(q[2])};
                const float target[3] = {float(q[0] + n[0]), float(q[1] + n[1]), float(q[2] + n[2])};
                const float size[2] = {float(sz[0]), float(sz[1])};
                check(tr_scene_add_rect_light(s, corner, target, size, float(I[0])), "tr_scene_add_rect_light");
            } else {
                throw std::runtime_error("GpuTransientRenderer: only point and rectangular area lights are rendered");
            }
        }
        if (Medium<3>* m = w.get_medium()) {
            auto* hm = static_cast<HomogeneousMedium<3, PhaseFunction::Isotropic<3>>*>(m);
            const Vector3 a = hm->get_absorption(VectorN<3>(0.)).to_rgb(), b = hm->get_scattering(VectorN<3>(0.)).to_rgb();
            const float sa[3] = {float(a[0]), float(a[1]), float(a[2])}, ss[3] = {float(b[0]), float(b[1]), float(b[2])};
            check(tr_scene_set_medium(s, sa, ss), "tr_scene_set_medium");
        }
        check(tr_scene_commit(s), "tr_scene_commit");
    } catch (...) {
        tr_scene_destroy(s);
        throw;
    }
    return s;
}

template <class Radiance, class RadianceAttenuation>
class GpuTransientRenderer : public RenderEngine<3, Radiance, RadianceAttenuation> {
    using Base = RenderEngine<3, Radiance, RadianceAttenuation>;

public:
    // camera and desc: what main would give PinholeCamera, StratifiedSampler and
    // BidirectionalPathTracing; scanline: -film-scanline (or -1)
    GpuTransientRenderer(FILE* f_log, std::vector<MeshEntry> meshes, const tr_camera& camera,
                         const tr_render_desc& desc, int scanline = -1)
        : Base(f_log), meshes_(std::move(meshes)), camera_(camera), desc_(desc), scanline_(scanline)
    {
    }

    void render(const char* name, World<3, Radiance>& world, Integrator<3, Radiance, RadianceAttenuation>* integrator,
                const Camera<3>& /*camera: settings come from camera_*/, Film<3, Radiance>* film,
                Sampler* /*sampler: settings come from desc_*/) const override
    {
        if (!integrator) throw std::runtime_error("Error: Need to define an integrator for rendering\n");
        tr_scene s = make_scene(world, meshes_);
        tr_film f = nullptr;
        try {
            const int W = film->get_width(), H = film->get_height(), T = int(film->get_time_resolution());
            tr_film_desc fd{W, H, T, float(film->get_exposure_time()), float(film->get_time_offset()), scanline_, 0, 0};
            check(tr_film_create(&f, &fd), "tr_film_create");
            tr_render_stats st{};
            check(tr_render_transient(s, &camera_, f, &desc_, &st), "tr_render_transient");
            if (this->f_log) std::fprintf(this->f_log, "Total time: GPU %f ms\n", st.render_ms);
            // back into the reference's film, scanline by scanline, through its own add_samples
            const int rows = scanline_ >= 0 ? 1 : H;
            std::vector<float> row(size_t(W) * size_t(T) * 3);
            for (int r = 0; r < rows; ++r) {
                check(tr_film_read_rows(f, r, 1, row.data(), nullptr), "tr_film_read_rows");
                const int y = scanline_ >= 0 ? scanline_ : r;
                for (int x = 0; x < W; ++x) {
                    RadianceSampleRecordVector<3, Radiance> rec;
                    for (int t = 0; t < T; ++t) {
                        const float* v = &row[(size_t(x) * size_t(T) + size_t(t)) * 3];
                        if (v[0] == 0.f && v[1] == 0.f && v[2] == 0.f) continue;
                        const Real time = film->get_time_offset() + (Real(t) + Real(0.5)) * film->get_exposure_time();
                        rec.add_sample(RadianceSample<3, Radiance>(Radiance(v[0], v[1], v[2]), time));
                    }
                    film->add_samples(Sample(Vector2(Real(x) + Real(0.5), Real(y) + Real(0.5)), Real(1.)), rec);
                }
            }
        } catch (...) {
            tr_film_destroy(f);
            tr_scene_destroy(s);
            throw;
        }
        tr_film_destroy(f);
        tr_scene_destroy(s);
        film->write(name);
    }

private:
    std::vector<MeshEntry> meshes_;
    tr_camera camera_;
    tr_render_desc desc_;
    int scanline_;
};

}  // namespace tr_reference

#endif  // TR_REFERENCE_ENGINE_HPP
