// Reference harness: drives the reference renderer's own classes (header-only,
// included from the reference source tree) in the reference's non-Embree
// configuration, so that golden fixtures can be produced from the reference's
// code without its CMake/Embree build.
//
// This file is TEST INFRASTRUCTURE. It is compiled only in a container where
// the reference sources exist (oracle/build_ref.sh) and its outputs are written
// as fixtures under tests/golden/. No product code links or calls it.
//
// What it exercises, by reference file:
//   TransientRenderer::render           src/Renderer/TransientRenderer.h:77-201
//   BidirectionalPathTracing (PT+NEE)   src/Integrator/BidirectionalPathTracing.h:486-692, 976-1140
//   StreakCameraFilm                    src/Film/StreakCameraFilm.h:139-192, 280-297, 357-437
//   StratifiedSampler, PinholeCamera    src/Sampling/StratifiedSampler.h:47-89, src/Camera/PinholeCamera.h:56-108
//   UberMaterial + Lambertian           src/Material/UberMaterial.h, src/Material/Reflectance/Lambertian.h
//   PointLightSource / Rectangular      src/LightSource/*.h
//   HomogeneousMedium + Isotropic       src/Media/*.h
//   Triangle::intersect (non-Embree)    src/Geometry/Primitives/3D/Triangle.h:112-163
//
// The reference hard-wires `#define _USE_EMBREE_` in bunnykiller.h. Including
// that header first and then undefining the macro selects the reference's own
// non-Embree World (World.h:665-708), which intersects through the reference's
// Triangle class instead of Embree. Triangles are handed to Triangle() with
// vertices 1 and 2 swapped so that its normal, normalize(cross(v1-v0, v2-v0)),
// equals the Embree-path convention normalize(cross(p2-p0, p1-p0)) (World.h:449).
//
// Input: a small text scene file (one directive per line, see parse_scene).
// Modes:
//   ref_harness render <scene.txt>          -> film slices written by the reference's
//                                              own StreakCameraFilm as .raw FP32 files
//   ref_harness render_cs <scene.txt>       -> same, global RNG reseeded per camera sample
//   ref_harness intersect <scene.txt> <rays.f32> <hits.bin>
//                                           -> brute-force closest hit with the reference's
//                                              Triangle::intersect: int32 id, float32 t per ray

#include "bunnykiller.h"
#undef _USE_EMBREE_

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <map>
#include <fstream>
#include <list>
#include <array>
#include <limits>
#include <sstream>

#include "Utils/Timer.h"
#include "Camera/PinholeCamera.h"
#include "RayTracing/World.h"
#include "Renderer/TransientRenderer.h"
#include "Film/StreakCameraFilm.h"
#include "Material/TwoSided.h"
#include "Film/MultibounceStreakCameraFilm.h"
#include "Filter/BoxFilter.h"
#include "Geometry/Aggregate/Mesh.h"
#include "Geometry/ObjectPointer.h"
#include "Material/UberMaterial.h"
#include "Material/Reflectance/Lambertian.h"
#include "LightSource/PointLightSource.h"
#include "LightSource/RectangularAreaLightSource.h"
#include "Sampling/StratifiedSampler.h"
#include "Integrator/BidirectionalPathTracing.h"
#include "Media/HomogeneousMedium.h"
#include "Media/PhaseFunction/IsotropicPhaseFunction.h"
#include "Utils/RandomNumbers.h"

typedef Spectrum Radiance;
typedef Spectrum RadianceAttenuation;

struct SceneDesc {
    int W = 16, H = 16, T = 256;
    double exposure = 0.01, offset = 0.0;
    int scanline = -1, unwarp = 0;
    double cam[10] = {0.5, 0.5, -1.2, 0.5, 0.5, 0.5, 0, 1, 0, 45.0};
    int sqrt_spp = 2, n_incoming = 1, max_bounces = 4;
    unsigned long long seed = 0;
    std::vector<std::array<double, 3>> materials;
    std::vector<std::array<double, 9>> tris;
    std::vector<int> tri_mat;
    std::vector<std::array<double, 9>> tri_vn;  // corner normals ("trin"); all zero = flat
    std::vector<int> two_sided;                  // material indices with TwoSidedMaterial
    std::vector<std::array<double, 6>> points;
    std::vector<std::array<double, 9>> rects;
    bool has_medium = false;
    double medium[6] = {0, 0, 0, 0, 0, 0};
    int multibounce = 0;  // > 0: MultibounceStreakFilm with this many bounce films (scanline only)
    bool time_sampling = false;  // -time-sampling: the integrator's Transient mode (bunnykiller.cpp:1811-1817)
    std::string out = "film";
};

static SceneDesc parse_scene(const char* path)
{
    SceneDesc s;
    std::ifstream in(path);
    if (!in) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#') continue;
        if (key == "film") {
            ls >> s.W >> s.H >> s.T >> s.exposure >> s.offset >> s.scanline >> s.unwarp;
        } else if (key == "camera") {
            for (double& v : s.cam) ls >> v;
        } else if (key == "render") {
            ls >> s.sqrt_spp >> s.n_incoming >> s.max_bounces >> s.seed;
        } else if (key == "material") {
            std::array<double, 3> m; ls >> m[0] >> m[1] >> m[2]; s.materials.push_back(m);
        } else if (key == "tri") {
            std::array<double, 9> t; for (double& v : t) ls >> v;
            int m; ls >> m; s.tris.push_back(t); s.tri_mat.push_back(m);
            s.tri_vn.push_back(std::array<double, 9>{});
        } else if (key == "trin") {
            std::array<double, 9> t, n; for (double& v : t) ls >> v; for (double& v : n) ls >> v;
            int m; ls >> m; s.tris.push_back(t); s.tri_mat.push_back(m); s.tri_vn.push_back(n);
        } else if (key == "twosided") {
            int m; ls >> m; s.two_sided.push_back(m);
        } else if (key == "point") {
            std::array<double, 6> p; for (double& v : p) ls >> v; s.points.push_back(p);
        } else if (key == "rect") {
            std::array<double, 9> r; for (double& v : r) ls >> v; s.rects.push_back(r);
        } else if (key == "medium") {
            s.has_medium = true; for (double& v : s.medium) ls >> v;
        } else if (key == "multibounce") {
            ls >> s.multibounce;
        } else if (key == "timesampling") {
            s.time_sampling = true;
        } else if (key == "out") {
            ls >> s.out;
        } else {
            fprintf(stderr, "unknown directive %s\n", key.c_str()); exit(2);
        }
    }
    return s;
}

static std::vector<Triangle> make_triangles(const SceneDesc& s, std::vector<Material3D*>& mats)
{
    std::vector<Triangle> tris;
    for (size_t i = 0; i < s.tris.size(); ++i) {
        const auto& t = s.tris[i];
        Vector3 v0(t[0], t[1], t[2]), v1(t[3], t[4], t[5]), v2(t[6], t[7], t[8]);
        const auto& q = s.tri_vn[i];
        bool smooth = false;
        for (double x : q) smooth = smooth || x != 0.0;
        // swapped: Triangle's normal = normalize(cross(b-a, c-a)) = normalize(cross(v2-v0, v1-v0));
        // corner normals follow their vertices
        if (smooth)
            tris.push_back(Triangle(v0, Vector3(q[0], q[1], q[2]), v2, Vector3(q[6], q[7], q[8]), v1,
                                    Vector3(q[3], q[4], q[5]), mats[s.tri_mat[i]]));
        else
            tris.push_back(Triangle(v0, v2, v1, mats[s.tri_mat[i]]));
    }
    return tris;
}

// Read access to the bounce films of a MultibounceStreakFilm and to protected members of a
// StreakCameraFilm, to dump the in-memory streak slice and weights of each bounce film.
using StreakR = StreakCameraFilm<3, Radiance>;
struct MBProbe : MultibounceStreakFilm<3, Radiance> {
    using MultibounceStreakFilm<3, Radiance>::MultibounceStreakFilm;
    StreakR* bounce(size_t b) { return m_bounce_films[b].get(); }
};
struct StreakProbe : StreakR {
    static Imaging::Image<Real> Film<3, Radiance>::*weight() { return &StreakProbe::m_weight; }
    static size_t StreakR::*first_slice() { return &StreakProbe::m_first_available_slice; }
};

static int run_render_mb(const SceneDesc& s, World<3, Radiance>& w, PinholeCamera& camera)
{
    if (s.scanline < 0) { fprintf(stderr, "multibounce dumps need a scanline film\n"); return 2; }
    StratifiedSampler sampler(s.W, s.H, s.sqrt_spp, true);
    MBProbe film(s.W, s.H, s.T, s.exposure, (unsigned)s.multibounce, new BoxFilter(1.), s.unwarp != 0,
                 FilmComponents::RADIANCE);
    film.set_name(s.out.c_str(), "raw");
    if (s.offset > 0.0) film.set_offset(s.offset);
    film.set_aspect_ratio(1.);
    sampler.set_scanline(s.scanline);
    film.set_streak(s.scanline);
    Random::StdRNG.seed(s.seed);
    FILE* f_log = fopen((s.out + ".log").c_str(), "w+");
    BidirectionalPathTracing<3, Radiance, RadianceAttenuation> bpt(w, s.n_incoming, f_log, &film);
    bpt.set_max_path_size(s.max_bounces);
    bpt.set_path_tracing_only(true);
    bpt.set_scattering_components(Scattering::ALL);
    bpt.set_mode(s.time_sampling ? BidirectionalPathTracing<3, Radiance, RadianceAttenuation>::SamplingMode::Transient
                                 : BidirectionalPathTracing<3, Radiance, RadianceAttenuation>::SamplingMode::SteadyState);
    TransientRenderer<3, Radiance, RadianceAttenuation> engine(f_log, s.time_sampling);
    engine.render(s.out.c_str(), w, &bpt, camera, &film, &sampler);
    if (f_log) fclose(f_log);
    // per bounce film: the unnormalised slice [W][T][3] and the weights [W] of the scanline
    FILE* fs = fopen((s.out + "_mb.f32").c_str(), "wb");
    FILE* fw = fopen((s.out + "_mbw.f32").c_str(), "wb");
    for (int b = 0; b < s.multibounce; ++b) {
        StreakR* f = film.bounce((size_t)b);
        const bool has_slice = (f->*StreakProbe::first_slice()) == (size_t)s.scanline;
        for (int x = 0; x < s.W; ++x)
            for (int t = 0; t < s.T; ++t)
                for (int c = 0; c < 3; ++c) {
                    float v = has_slice ? (float)f->streak_image(s.scanline)(t, x, c) : 0.f;
                    fwrite(&v, 4, 1, fs);
                }
        const Imaging::Image<Real>& wimg = f->*StreakProbe::weight();
        for (int x = 0; x < s.W; ++x) {
            float v = (float)wimg(x, s.scanline, 0);
            fwrite(&v, 4, 1, fw);
        }
    }
    fclose(fs);
    fclose(fw);
    return 0;
}

static int run_render(const SceneDesc& s, bool reseed_per_camera_sample)
{
    World<3, Radiance> w;
    std::vector<Material3D*> mats;
    for (size_t mi = 0; mi < s.materials.size(); ++mi) {
        const auto& m = s.materials[mi];
        UberMaterial3D* um = new UberMaterial3D();
        um->add_bsdf(new Lambertian3D(Spectrum(m[0], m[1], m[2])));
        bool two = std::find(s.two_sided.begin(), s.two_sided.end(), (int)mi) != s.two_sided.end();
        mats.push_back(two ? static_cast<Material3D*>(new TwoSidedMaterial3D(um)) : static_cast<Material3D*>(um));
    }
    // one Mesh per material, triangles in scene order
    for (size_t mi = 0; mi < mats.size(); ++mi) {
        std::vector<Triangle> group;
        std::vector<Triangle> all = make_triangles(s, mats);
        for (size_t i = 0; i < all.size(); ++i)
            if (s.tri_mat[i] == (int)mi) group.push_back(all[i]);
        if (!group.empty())
            w.add_object(Object3DPointer(new Mesh(group, static_cast<Material3D*>(mats[mi]))));
    }
    for (const auto& p : s.points)
        w.add_light(new PointLightSource<3, Radiance>(&w, Vector3(p[0], p[1], p[2]),
                                                      Radiance(p[3], p[4], p[5])));
    for (const auto& r : s.rects) {
        Vector3 pos(r[0], r[1], r[2]), tgt(r[3], r[4], r[5]);
        w.add_light(new RectangularAreaLightSource<Radiance>(&w, pos, (tgt - pos).normalized(),
                                                             Vector2(r[6], r[7]), Radiance(r[8])));
    }
    if (s.has_medium) {
        w.set_medium(new HomogeneousMedium<3, PhaseFunction::Isotropic<3>>(
            Spectrum(s.medium[0], s.medium[1], s.medium[2]),
            Spectrum(s.medium[3], s.medium[4], s.medium[5]),
            PhaseFunction::Isotropic<3>(0.)));
    }

    // -camera-fov: view_plane_dist = 1/tan(fov/2) (bunnykiller.cpp:1156-1160)
    Real fov = M_PI / 180. * s.cam[9] / 2;
    PinholeCamera camera(Vector3(s.cam[0], s.cam[1], s.cam[2]), Vector3(s.cam[3], s.cam[4], s.cam[5]),
                         Vector3(s.cam[6], s.cam[7], s.cam[8]), 1. / tan(fov));

    w.set_background(Spectrum(0.));
    w.freeze();
    if (s.multibounce > 0) return run_render_mb(s, w, camera);

    StratifiedSampler sampler(s.W, s.H, s.sqrt_spp, true);
    StreakCameraFilm<3, Radiance> film(s.W, s.H, s.T, s.exposure, s.unwarp != 0, new BoxFilter(1.),
                                       FilmComponents::RADIANCE);
    film.set_name(s.out.c_str(), "raw");
    if (s.offset > 0.0) film.set_offset(s.offset);
    film.set_aspect_ratio(1.);
    if (s.scanline > -1) {
        sampler.set_scanline(s.scanline);
        film.set_streak(s.scanline);
    }

    Random::StdRNG.seed(s.seed);

    FILE* f_log = fopen((s.out + ".log").c_str(), "w+");
    BidirectionalPathTracing<3, Radiance, RadianceAttenuation> bpt(w, s.n_incoming, f_log, &film);
    bpt.set_max_path_size(s.max_bounces);
    bpt.set_path_tracing_only(true);
    bpt.set_scattering_components(Scattering::ALL);
    bpt.set_mode(s.time_sampling ? BidirectionalPathTracing<3, Radiance, RadianceAttenuation>::SamplingMode::Transient
                                 : BidirectionalPathTracing<3, Radiance, RadianceAttenuation>::SamplingMode::SteadyState);

    if (!reseed_per_camera_sample) {
        TransientRenderer<3, Radiance, RadianceAttenuation> engine(f_log, s.time_sampling);
        engine.render(s.out.c_str(), w, &bpt, camera, &film, &sampler);
    } else {
        // The loop of TransientRenderer::render (TransientRenderer.h:130-179) with the global RNG
        // reseeded before each camera sample as Random::StdRNG.seed(seed, cs), cs being the
        // camera sample's index ((y*W + x)*spp^2 + sub-sample), so that any divergence between two
        // implementations stays within one camera sample.
        Sample fs;
        RadianceSampleRecordVector<3, Radiance> rec;
        unsigned long long spp2 = (unsigned long long)s.sqrt_spp * s.sqrt_spp;
        unsigned long long cs = (s.scanline > -1) ? (unsigned long long)s.scanline * s.W * spp2 : 0ULL;
        for (;; ++cs) {
            Random::StdRNG.seed(s.seed, cs);
            if (!sampler.get_next_sample(fs)) break;
            Vector2 meh = film.window_coords2image_coords(fs.position);
            VectorN<2> ic;
            ic[0] = meh[0];
            ic[1] = meh[1];
            Ray<3> r = camera.get_ray(ic);
            if (s.time_sampling) w.Li(r, &bpt, rec, film.get_time_length(), film.get_time_resolution());
            else w.Li(r, &bpt, rec, film.get_exposure_time(), 0);
            for (auto& smp : rec.samples)
                if (!smp.radiance.is_finite()) smp.radiance = Radiance(0.);
            film.add_samples(fs, rec);
            rec.samples.clear();
        }
        film.write(s.out.c_str());
    }
    if (f_log) fclose(f_log);
    return 0;
}

static int run_intersect(const SceneDesc& s, const char* rays_path, const char* hits_path)
{
    std::vector<Material3D*> mats;
    for (size_t i = 0; i < s.materials.size(); ++i) {
        UberMaterial3D* um = new UberMaterial3D();
        um->add_bsdf(new Lambertian3D(Spectrum(.5)));
        mats.push_back(um);
    }
    std::vector<Triangle> tris = make_triangles(s, mats);
    FILE* fr = fopen(rays_path, "rb");
    if (!fr) { fprintf(stderr, "cannot open %s\n", rays_path); return 2; }
    std::vector<float> rays;
    float buf[6];
    while (fread(buf, sizeof(float), 6, fr) == 6) rays.insert(rays.end(), buf, buf + 6);
    fclose(fr);
    size_t n = rays.size() / 6;
    FILE* fh = fopen(hits_path, "wb");
    for (size_t r = 0; r < n; ++r) {
        const float* q = &rays[r * 6];
        Ray3D ray(Vector3(q[0], q[1], q[2]), Vector3(q[3], q[4], q[5]), true);
        Intersection3D it;
        int32_t id = -1;
        for (size_t i = 0; i < tris.size(); ++i)
            if (tris[i].intersect(ray, it, std::numeric_limits<float>::infinity()))
                id = (int32_t)i;
        float t = (id >= 0) ? (float)ray.get_parameter() : -1.f;
        fwrite(&id, 4, 1, fh);
        fwrite(&t, 4, 1, fh);
    }
    fclose(fh);
    return 0;
}

// The reference's global RNG after Random::StdRNG.seed(state[, seq]): n integers and n floats.
static int run_rng(int argc, char** argv)
{
    unsigned long long state = strtoull(argv[2], nullptr, 0);
    int n = atoi(argv[3]);
    if (argc >= 5) Random::StdRNG.seed(state, strtoull(argv[4], nullptr, 0));
    else Random::StdRNG.seed(state);
    Random::RNG copy(Random::StdRNG);
    for (int i = 0; i < n; ++i) printf("%u%c", Random::StdRNG.next_int(), i + 1 < n ? ' ' : '\n');
    for (int i = 0; i < n; ++i) printf("%.9g%c", copy.next_float(), i + 1 < n ? ' ' : '\n');
    return 0;
}

// save <in.f32> <W> <H> <out.hdr|out.raw>: an RGB float image through Imaging::save (ImageIO.h)
static int run_save(int argc, char** argv)
{
    (void)argc;
    unsigned W = unsigned(atoi(argv[3])), H = unsigned(atoi(argv[4]));
    std::vector<float> px(size_t(W) * H * 3);
    FILE* f = fopen(argv[2], "rb");
    if (!f || fread(px.data(), sizeof(float), px.size(), f) != px.size()) {
        fprintf(stderr, "cannot read %s\n", argv[2]);
        return 1;
    }
    fclose(f);
    std::unique_ptr<Real[]> buf(new Real[px.size()]);
    std::copy(px.begin(), px.end(), buf.get());
    Imaging::Image<Real> img(std::move(buf), W, H, 3);
    Imaging::save(img, argv[5]);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc >= 6 && !strcmp(argv[1], "save")) return run_save(argc, argv);
    if (argc >= 4 && !strcmp(argv[1], "rng")) return run_rng(argc, argv);
    if (argc >= 3 && !strcmp(argv[1], "render")) return run_render(parse_scene(argv[2]), false);
    if (argc >= 3 && !strcmp(argv[1], "render_cs")) return run_render(parse_scene(argv[2]), true);
    if (argc >= 5 && !strcmp(argv[1], "intersect"))
        return run_intersect(parse_scene(argv[2]), argv[3], argv[4]);
    fprintf(stderr, "usage: ref_harness render|render_cs <scene> | intersect <scene> <rays.f32> <hits.bin>\n");
    return 2;
}
