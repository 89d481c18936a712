// tr_bunnykiller.hpp — C++ host classes over the C ABI (tr_api.h) that keep the reference
// renderer's interfaces for the transient path, so that its driver code ports one-for-one:
//
//   reference (bunnykiller.cpp:1504-1827)            this header
//   World<3,Radiance> w; w.add_light(...)            tr::World w; w.add_light(...)
//   w.set_medium(new HomogeneousMedium<..>(a,s,pf))  w.set_medium(tr::HomogeneousMedium(a, s))
//   w.freeze()                                       w.freeze()
//   PinholeCamera(c, f, up, view_plane_dist)         tr::PinholeCamera(c, f, up, view_plane_dist)
//   StratifiedSampler(W, H, sqrt_spp, true)          tr::StratifiedSampler(W, H, sqrt_spp, true)
//   StreakCameraFilm(W, H, T, exp, unwarp, filter)   tr::StreakCameraFilm(W, H, T, exp, unwarp)
//   BidirectionalPathTracing bpt(w, N, log, film)    tr::BidirectionalPathTracing bpt(w, N, log, &film)
//   bpt.set_max_path_size / set_path_tracing_only    same
//   TransientRenderer(f_log).render(name, w, &bpt,   tr::TransientRenderer(f_log).render(name, w, &bpt,
//                      camera, film, &sampler)                          camera, &film, &sampler)
//
// Errors are reported as std::runtime_error, as the reference does (TransientRenderer.h:86-88,
// World.h:211-214). Every render runs on the GPU through tr_render_transient.
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tr_api.h"

namespace tr {

using Real = float;
using Vector3 = std::array<Real, 3>;

inline void check(int rc, const char* what)
{
    if (rc != TR_OK) throw std::runtime_error(std::string(what) + ": " + tr_last_error());
}

// Lambertian BSDF in a one-BSDF UberMaterial (-lambertian r g b, bunnykiller.cpp:328-336)
struct Lambertian {
    Vector3 Kd;
    explicit Lambertian(Vector3 kd) : Kd(kd) {}
};

// PointLightSource (src/LightSource/PointLightSource.h:39-171)
struct PointLightSource {
    Vector3 position, intensities;
    PointLightSource(Vector3 pos, Vector3 ints) : position(pos), intensities(ints) {}
};

// RectangularAreaLightSource as built by -rectangular-area-light-source (bunnykiller.cpp:1242-1261):
// corner, target point (normal = normalize(target - corner)), size, scalar intensity
struct RectangularAreaLightSource {
    Vector3 position, target;
    std::array<Real, 2> size;
    Real intensity;
    RectangularAreaLightSource(Vector3 pos, Vector3 tgt, std::array<Real, 2> sz, Real ints)
        : position(pos), target(tgt), size(sz), intensity(ints)
    {
    }
};

// HomogeneousMedium<Isotropic> without phase time delay (src/Media/HomogeneousMedium.h:45-127)
struct HomogeneousMedium {
    Vector3 absorption, scattering;
    HomogeneousMedium(Vector3 a, Vector3 s) : absorption(a), scattering(s) {}
};

// World (src/RayTracing/World.h:95-789)
class World {
public:
    World() { check(tr_scene_create(&h_), "World"); }
    ~World() { tr_scene_destroy(h_); }
    World(const World&) = delete;
    World& operator=(const World&) = delete;

    // World::add_triangle_mesh (World.h:209-541) from arrays: xyz (3 per vertex), idx (3 per triangle)
    unsigned add_triangle_mesh(const std::vector<float>& xyz, const std::vector<uint32_t>& idx, const Lambertian& m)
    {
        if (xyz.size() % 3 || idx.size() % 3) throw std::runtime_error("add_triangle_mesh: arrays not multiples of 3");
        tr_material mat{{m.Kd[0], m.Kd[1], m.Kd[2]}};
        uint32_t g = 0;
        check(tr_scene_add_mesh(h_, xyz.data(), uint32_t(xyz.size() / 3), idx.data(), uint32_t(idx.size() / 3), &mat,
                                &g),
              "add_triangle_mesh");
        frozen_ = false;
        return g;
    }
    // World::add_triangle_mesh(name_file, mat) (World.h:209-541): a Wavefront OBJ file; a
    // two-sided material is TwoSidedMaterial around the Lambertian (-two-sided, bunnykiller.cpp:408-412)
    unsigned add_triangle_mesh(const std::string& obj_file, const Lambertian& m, bool two_sided = false)
    {
        tr_material mat{{m.Kd[0], m.Kd[1], m.Kd[2]}};
        uint32_t g = 0;
        check(tr_scene_add_obj(h_, obj_file.c_str(), &mat, two_sided ? TR_MATERIAL_TWO_SIDED : 0u, &g),
              "add_triangle_mesh");
        frozen_ = false;
        return g;
    }
    void add_light(const PointLightSource& l)
    {
        check(tr_scene_add_point_light(h_, l.position.data(), l.intensities.data()), "add_light");
        ++lights_;
        frozen_ = false;
    }
    void add_light(const RectangularAreaLightSource& l)
    {
        check(tr_scene_add_rect_light(h_, l.position.data(), l.target.data(), l.size.data(), l.intensity), "add_light");
        ++lights_;
        frozen_ = false;
    }
    unsigned nb_lights() const { return lights_; }
    void set_medium(const HomogeneousMedium& m)
    {
        check(tr_scene_set_medium(h_, m.absorption.data(), m.scattering.data()), "set_medium");
        frozen_ = false;
    }
    // World::freeze (World.h:543-550)
    void freeze()
    {
        check(tr_scene_commit(h_), "freeze");
        frozen_ = true;
    }
    bool frozen() const { return frozen_; }
    // batched World::first_intersection (World.h:566-615)
    std::vector<tr_hit> first_intersection(const std::vector<tr_ray>& rays) const
    {
        std::vector<tr_hit> hits(rays.size());
        check(tr_intersect_batch(h_, rays.data(), uint32_t(rays.size()), hits.data()), "first_intersection");
        return hits;
    }
    // batched World::intersects (World.h:617-639): 1 where a triangle is hit in (tnear, tfar]
    std::vector<uint8_t> intersects(const std::vector<tr_ray>& rays) const
    {
        std::vector<uint8_t> occ(rays.size());
        check(tr_occluded_batch(h_, rays.data(), uint32_t(rays.size()), occ.data()), "intersects");
        return occ;
    }
    tr_scene handle() const { return h_; }

private:
    tr_scene h_ = nullptr;
    unsigned lights_ = 0;
    bool frozen_ = false;
};

// PinholeCamera (src/Camera/PinholeCamera.h:56-108): fd is the view-plane distance
// (-camera-fov f gives fd = 1/tan(f/2), bunnykiller.cpp:1156-1160)
class PinholeCamera {
public:
    PinholeCamera(Vector3 c, Vector3 f, Vector3 u, Real fd) : c_(c), f_(f), u_(u), fd_(fd) {}
    static PinholeCamera from_fov(Vector3 c, Vector3 f, Vector3 u, Real fov_deg)
    {
        Real fov = Real(M_PI / 180. * fov_deg / 2);
        return PinholeCamera(c, f, u, Real(1. / std::tan(fov)));
    }
    tr_camera desc() const
    {
        tr_camera d{};
        for (int k = 0; k < 3; ++k) {
            d.position[k] = c_[k];
            d.focus[k] = f_[k];
            d.up[k] = u_[k];
        }
        d.fov_deg = Real(2.0 * std::atan(1.0 / double(fd_)) * 180.0 / M_PI);
        d.view_plane_dist = fd_;
        return d;
    }

private:
    Vector3 c_, f_, u_;
    Real fd_;
};

// StratifiedSampler (src/Sampling/StratifiedSampler.h:32-125)
class StratifiedSampler {
public:
    StratifiedSampler(int size_x, int size_y, int sqrt_spp, bool jittered = true)
        : x_(size_x), y_(size_y), spp_(sqrt_spp)
    {
        if (!jittered) throw std::runtime_error("StratifiedSampler: only the jittered sampler is provided");
    }
    void set_scanline(int y) { scanline_ = y; }
    int scanline() const { return scanline_; }
    int sqrt_spp() const { return spp_; }
    int size_x() const { return x_; }
    int size_y() const { return y_; }
    size_t get_nb_samples() const { return size_t(scanline_ >= 0 ? 1 : y_) * x_ * spp_ * spp_; }

private:
    int x_, y_, spp_, scanline_ = -1;
};

// StreakCameraFilm (src/Film/StreakCameraFilm.h:40-437), box filter of size 1
class StreakCameraFilm {
public:
    StreakCameraFilm(unsigned w, unsigned h, unsigned t, Real exposure, bool camera_unwarp = false)
    {
        d_ = tr_film_desc{int32_t(w), int32_t(h), int32_t(t), exposure, 0.f, -1, camera_unwarp ? 1 : 0, 0};
    }
    ~StreakCameraFilm()
    {
        if (h_) tr_film_destroy(h_);
    }
    StreakCameraFilm(const StreakCameraFilm&) = delete;
    StreakCameraFilm& operator=(const StreakCameraFilm&) = delete;

    void set_offset(Real offset) { unallocated(), d_.offset = offset; }
    void set_streak(unsigned y) { unallocated(), d_.scanline = int32_t(y); }
    void set_name(const std::string& name, const std::string& ext = "raw") { name_ = name, ext_ = ext; }
    Real get_time_length() const { return Real(d_.time_bins) * d_.exposure + d_.offset; }
    Real get_exposure_time() const { return d_.exposure; }
    Real get_time_offset() const { return d_.offset; }
    unsigned get_time_resolution() const { return unsigned(d_.time_bins); }
    unsigned get_width() const { return unsigned(d_.width); }
    unsigned get_height() const { return unsigned(d_.height); }
    int rows() const { return d_.scanline >= 0 ? 1 : d_.height; }

    tr_film handle()
    {
        if (!h_) check(tr_film_create(&h_, &d_), "StreakCameraFilm");
        return h_;
    }
    void clean() { check(tr_film_clear(handle()), "StreakCameraFilm::clean"); }
    // normalised streak data: [row][x][t][rgb]
    std::vector<float> data()
    {
        std::vector<float> out(size_t(d_.bounce_layers > 0 ? d_.bounce_layers : 1) * size_t(rows()) * d_.width *
                               d_.time_bins * 3);
        check(tr_film_read(handle(), out.data(), nullptr), "StreakCameraFilm::data");
        return out;
    }
    // StreakCameraFilm::write + Film::write (StreakCameraFilm.h:233-242, 357-437; Film.h:358-471):
    // "<name>_%04d.<ext>" per scanline (T wide, W high; transposed when T < 8), ext "hdr" (RGBE)
    // or "raw" (FP32), then the steady image "<name>.hdr"
    void write(const char* name = nullptr)
    {
        std::string base = name ? name : name_;
        if (base.empty()) throw std::runtime_error("StreakCameraFilm::write: no name");
        check(tr_film_write(handle(), base.c_str(), ext_.c_str()), "StreakCameraFilm::write");
    }

protected:
    void unallocated() const
    {
        if (h_) throw std::runtime_error("StreakCameraFilm: geometry is fixed once rendering started");
    }
    tr_film_desc d_{};
    tr_film h_ = nullptr;
    std::string name_, ext_ = "raw";
};

// MultibounceStreakFilm (src/Film/MultibounceStreakCameraFilm.h:67-171): one streak film per
// bounce ("<name>_bXX_%04d.<ext>", XX = bounce + 1; bounces >= nb_bounces go to the last one)
class MultibounceStreakFilm : public StreakCameraFilm {
public:
    MultibounceStreakFilm(unsigned w, unsigned h, unsigned t, Real exposure, unsigned nb_bounces,
                          bool camera_unwarp = false)
        : StreakCameraFilm(w, h, t, exposure, camera_unwarp)
    {
        if (nb_bounces == 0 || nb_bounces > 32) throw std::runtime_error("MultibounceStreakFilm: 1..32 bounces");
        d_.bounce_layers = int32_t(nb_bounces);
    }
    unsigned nb_bounces() const { return unsigned(d_.bounce_layers); }
};

// BidirectionalPathTracing in path-tracing-only mode (BidirectionalPathTracing.h:122-217) with the
// SteadyState sampling techniques or the Transient ones (set_mode, :190-207)
class BidirectionalPathTracing {
public:
    enum SamplingMode { SteadyState, Transient };
    BidirectionalPathTracing(World& w, unsigned incoming_samples = 1, FILE* f_log = nullptr,
                             StreakCameraFilm* film = nullptr)
        : w_(w), n_(incoming_samples), log_(f_log), film_(film)
    {
        if (incoming_samples == 0) throw std::runtime_error("BidirectionalPathTracing: zero incoming samples");
    }
    void set_max_path_size(int n) { max_ = n; }
    void set_path_tracing_only(bool on)
    {
        if (!on) throw std::runtime_error("BidirectionalPathTracing: light subpaths are not provided");
    }
    void set_mode(SamplingMode m) { mode_ = m; }
    SamplingMode mode() const { return mode_; }
    void set_seed(uint64_t s) { seed_ = s; }
    void preprocess() const
    {
        if (!log_) return;
        const bool t = mode_ == Transient;
        std::fprintf(log_, "Bidirectional Path Tracing\n- Samples: %u\n- Next Event: %s\n", n_,
                     t ? "Time Sampling - #Vertices 4" : "Mean Free Path Sampling");
        std::fprintf(log_, "- Angular Sampling: %s\n- Path Termination (Max Path Size %d): %s\n"
                           "- Vertices Connection: %s\n\n",
                     t ? "Shortest Time" : "Phase Function", max_, t ? "Total Time" : "Russian Roulette",
                     t ? "Point-to-Line" : "Shadow Connection");
    }
    unsigned incoming_samples() const { return n_; }
    int max_path_size() const { return max_; }
    uint64_t seed() const { return seed_; }

private:
    World& w_;
    unsigned n_;
    FILE* log_;
    StreakCameraFilm* film_;
    int max_ = 30;  // constructor default (BidirectionalPathTracing.h:131)
    SamplingMode mode_ = SteadyState;
    uint64_t seed_ = 0x853c49e6748fea9bULL;  // Random::RNG default state, as without -new-seed
};

// TransientRenderer (src/Renderer/TransientRenderer.h:77-201)
class TransientRenderer {
public:
    // time_sampling only changes the time arguments World::Li passes on, which the path-tracing
    // integrator ignores (TransientRenderer.h:160-164); the integrator's mode selects the techniques
    explicit TransientRenderer(FILE* f_log = nullptr, bool time_sampling = false) : log_(f_log)
    {
        (void)time_sampling;
    }
    tr_render_stats render(const char* name, World& world, BidirectionalPathTracing* integrator,
                           const PinholeCamera& camera, StreakCameraFilm* film, StratifiedSampler* sampler) const
    {
        if (!integrator) throw std::runtime_error("Error: Need to define an integrator for rendering\n");
        if (!film || !sampler) throw std::runtime_error("TransientRenderer: film and sampler are required");
        if (log_) std::fprintf(log_, "RENDER\n======\n");
        integrator->preprocess();
        if (!world.frozen()) world.freeze();
        if (sampler->scanline() >= 0) film->set_streak(unsigned(sampler->scanline()));
        tr_camera cam = camera.desc();
        tr_render_desc d{};
        d.sqrt_spp = sampler->sqrt_spp();
        d.n_incoming = int32_t(integrator->incoming_samples());
        d.max_bounces = integrator->max_path_size();
        d.seed = integrator->seed();
        if (integrator->mode() == BidirectionalPathTracing::Transient) d.flags |= TR_RENDER_TIME_SAMPLING;
        tr_render_stats st{};
        check(tr_render_transient(world.handle(), &cam, film->handle(), &d, &st), "TransientRenderer::render");
        if (log_) {
            double secs = st.render_ms * 1e-3;
            unsigned hours = unsigned(secs) / 3600;
            secs -= hours * 3600;
            unsigned minutes = unsigned(secs) / 60;
            secs -= minutes * 60;
            std::fprintf(log_, "Total time: %u:%u:%f\n", hours, minutes, secs);
        }
        if (name) film->write(name);
        return st;
    }

private:
    FILE* log_;
};

}  // namespace tr
