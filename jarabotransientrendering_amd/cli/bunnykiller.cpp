// bunnykiller: the reference renderer's command line (bunnykiller.cpp:980-1502, main :1504-1827)
// for the transient path this library accelerates: the same flags build the same scene, film,
// camera and integrator, and the render runs on the GPU (tr_bunnykiller.hpp over the C ABI).
//
// Supported flags (defaults as ProgramParams, src/Utils/ProgramParams.h:115-168):
//   -film-name N  -film-size-x W  -film-size-y H  -film-size-t T  -film-exposure E  -film-offset O
//   -film-scanline Y  -camera-unwarp  -multibounce-streak B
//   -camera-position x y z  -camera-focus x y z  -camera-up x y z  -camera-view-plane d
//   -camera-fov deg  -camera-spp sqrt_spp
//   -point-light-source x y z r g b
//   -rectangular-area-light-source cx cy cz tx ty tz sx sy I
//   -name-mesh file.obj [-lambertian r g b | -two-sided -lambertian r g b]  -lambertian-rho g
//   -homogeneous-medium sa(3) ss(3) [-pf-isotropic] [-pf-time 0]
//   -max-nb-bounces N  -bidirectional-path-tracing N  -new-seed S  -log-name F
//   -transient-state  -standard-sampling  -time-sampling  -scattering-level all  -parse-args-file F
// Flags of features outside the transient path-tracing path (other integrators, lights, BSDFs,
// phase functions, steady-state films) stop with an error instead of being
// ignored; unknown arguments are skipped, as the reference does (bunnykiller.cpp:1491-1494).
// Output: "<name>_%04d.hdr" streak images and "<name>.hdr" (multibounce: "<name>_bXX_%04d.hdr").
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "tr_bunnykiller.hpp"

namespace {

struct MeshArg {
    std::string file;
    tr::Vector3 kd;
    bool two_sided;
    bool default_material;  // no material flag: the grey default, whose albedo is the last -lambertian-rho
};

struct Params {
    tr::Vector3 cam_pos{0.f, 0.f, 0.f}, cam_focus{1.f, 0.f, 0.f}, cam_up{0.f, 1.f, 0.f};
    float view_plane = 1.f;
    float grey = 0.3f;
    int width = 256, height = 256, time = 512, scanline = -1, sqrt_spp = 2;
    float exposure = 100.f, offset = 0.f;
    bool unwarp = false;
    int multibounce = 0;
    std::string image_file = "name_file", log_file;
    uint64_t seed = 0x853c49e6748fea9bULL;  // Random::RNG::default_state
    unsigned n_incoming = 1024;
    int max_bounces = 20;
    std::vector<MeshArg> meshes;
    std::vector<tr::PointLightSource> points;
    std::vector<tr::RectangularAreaLightSource> rects;
    bool has_medium = false;
    tr::Vector3 sa{}, ss{};
    bool time_sampling = false;
    bool print_parsed = false;  // --print-parsed: print what the command line set up, render nothing
};

[[noreturn]] void unsupported(const std::string& flag)
{
    throw std::runtime_error("Error: " + flag + " is outside the transient path-tracing path of this renderer");
}

struct Args {
    std::vector<std::string> a;
    size_t i = 0;
    bool more(size_t k = 1) const { return i + k <= a.size(); }
    const std::string& peek() const { return a[i]; }
    std::string next()
    {
        if (i >= a.size()) throw std::runtime_error("Error: missing argument after " + (i ? a[i - 1] : std::string()));
        return a[i++];
    }
    float f() { return float(std::atof(next().c_str())); }
    int n() { return std::atoi(next().c_str()); }
    tr::Vector3 v3()
    {
        float x = f(), y = f(), z = f();
        return {x, y, z};
    }
};

void parse(Args& A, Params& p);

// load_material (bunnykiller.cpp:323-416): -lambertian, -two-sided <material>; default grey
void load_material(Args& A, Params& p, MeshArg& m)
{
    if (!A.more()) return;
    if (A.peek() == "-lambertian") {
        A.next();
        m.kd = A.v3();
        m.default_material = false;
    } else if (A.peek() == "-two-sided") {
        A.next();
        m.two_sided = true;
        load_material(A, p, m);
    } else if (A.peek() == "-phong" || A.peek() == "-ward" || A.peek() == "-transparent" ||
               A.peek() == "-dielectric" || A.peek() == "-conductor" || A.peek() == "-lambertian-textured") {
        unsupported(A.peek());
    }
}

void parse_file(const std::string& path, Params& p)
{
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "WARNING: cannot read args file %s\n", path.c_str());
        return;
    }
    Args A;
    A.a.push_back("bunnykiller");
    std::string tok;
    while (in >> tok) A.a.push_back(tok);
    A.i = 1;
    parse(A, p);
}

void parse(Args& A, Params& p)
{
    while (A.i < A.a.size()) {
        const std::string f = A.next();
        if (f == "-lambertian-rho") p.grey = A.f();
        else if (f == "-name-mesh") {
            MeshArg m{A.next(), {p.grey, p.grey, p.grey}, false, true};
            load_material(A, p, m);
            p.meshes.push_back(m);
        } else if (f == "-film-name") p.image_file = A.next();
        else if (f == "-film-size-x") p.width = p.height = A.n();
        else if (f == "-film-size-y") p.height = A.n();
        else if (f == "-film-size-t") p.time = A.n();
        else if (f == "-film-exposure") p.exposure = A.f();
        else if (f == "-film-scanline") p.scanline = A.n();
        else if (f == "-film-offset") p.offset = A.f();
        else if (f == "-film-aspect-ratio") {
            if (A.f() != 1.f) unsupported("-film-aspect-ratio other than 1");
        } else if (f == "-camera-unwarp") p.unwarp = true;
        else if (f == "-multibounce-streak") p.multibounce = A.n();
        else if (f == "-camera-position") p.cam_pos = A.v3();
        else if (f == "-camera-focus") p.cam_focus = A.v3();
        else if (f == "-camera-up") {
            tr::Vector3 u = A.v3();
            float l = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            p.cam_up = {u[0] / l, u[1] / l, u[2] / l};
        } else if (f == "-camera-view-plane") p.view_plane = A.f();
        else if (f == "-camera-spp") p.sqrt_spp = A.n();
        else if (f == "-camera-fov") {
            float fov = float(M_PI / 180. * A.f() / 2);
            p.view_plane = float(1. / std::tan(fov));
        } else if (f == "-point-light-source") {
            tr::Vector3 pos = A.v3(), I = A.v3();
            p.points.emplace_back(pos, I);
        } else if (f == "-rectangular-area-light-source") {
            tr::Vector3 pos = A.v3(), tgt = A.v3();
            float sx = A.f(), sy = A.f(), I = A.f();
            // the reference steps over the next argument while probing for -emission-texture
            // (bunnykiller.cpp:1253): keep that, so that the same command lines parse alike
            if (A.a.size() > A.i + 2) {
                if (A.next() == "-emission-texture") unsupported("-emission-texture");
            }
            p.rects.emplace_back(pos, tgt, std::array<float, 2>{sx, sy}, I);
        } else if (f == "-homogeneous-medium") {
            p.sa = A.v3();
            p.ss = A.v3();
            p.has_medium = true;
            if (A.more() && A.peek() == "-pf-isotropic") A.next();
            else if (A.more() && (A.peek() == "-pf-hg" || A.peek() == "-pf-rayleigh")) unsupported(A.peek());
            if (A.a.size() > A.i + 2 && A.peek() == "-pf-time") {
                A.next();
                if (A.f() != 0.f) unsupported("-pf-time other than 0");
            }
        } else if (f == "-max-nb-bounces") p.max_bounces = A.n();
        else if (f == "-bidirectional-path-tracing") p.n_incoming = unsigned(A.n());
        else if (f == "-new-seed") p.seed = uint64_t(std::atoi(A.next().c_str()));  // atoi, as the reference
        else if (f == "-log-name") p.log_file = A.next();
        else if (f == "-transient-state") {
        } else if (f == "-standard-sampling") p.time_sampling = false;
        else if (f == "-time-sampling") p.time_sampling = true;
        else if (f == "-scattering-level") {
            // multiple / single / none select other components; anything else is ALL (bunnykiller.cpp:1372-1385)
            const std::string lv = A.next();
            if (lv == "multiple" || lv == "single" || lv == "none") unsupported("-scattering-level " + lv);
        } else if (f == "-parse-args-file") parse_file(A.next(), p);
        else if (f == "--print-parsed") p.print_parsed = true;
        else if (f == "-steady-state" || f == "-photon-mapping" || f == "-sphere" ||
                 f == "-plane" || f == "-mie-medium" || f == "-cosine-light-source" ||
                 f == "-hemispherical-light-source" || f == "-point-light-source-polarized" ||
                 f == "-spot-light-source" || f == "-directional-light-source" ||
                 f == "-virtual-point-light-emitter" || f == "-virtual-point-light-source" ||
                 f == "-kernel-based-streak" || f == "-film-single-pixel" || f == "-film-reveal" ||
                 f == "-film-store-depth" || f == "-film-store-positions" || f == "-film-store-normals" ||
                 f == "-parse-xml-file" || f == "-camera-view-plane-size")
            unsupported(f);
        // anything else: skipped, as the reference skips unmatched arguments
    }
}

void vec(const char* key, const tr::Vector3& v, bool comma = true)
{
    std::printf("\"%s\": [%.9g, %.9g, %.9g]%s", key, double(v[0]), double(v[1]), double(v[2]), comma ? ", " : "");
}

// What the command line set up, as JSON: the film, camera, sampler and integrator settings, the lights
// (with main's default light), the medium and every mesh with its material and triangle count.
int print_parsed(const Params& p)
{
    std::printf("{\"film\": {\"width\": %d, \"height\": %d, \"time_bins\": %d, \"exposure\": %.9g, "
                "\"offset\": %.9g, \"scanline\": %d, \"aspect_ratio\": 1, \"camera_unwarp\": %d, "
                "\"bounce_layers\": %d, \"transient\": 1, \"image_file\": \"%s\", \"log_file\": \"%s\"}, ",
                p.width, p.height, p.time, double(p.exposure), double(p.offset), p.scanline, int(p.unwarp), p.multibounce,
                p.image_file.c_str(), p.log_file.c_str());
    std::printf("\"camera\": {");
    vec("position", p.cam_pos);
    vec("focus", p.cam_focus);
    vec("up", p.cam_up);
    std::printf("\"view_plane_dist\": %.9g}, ", double(p.view_plane));
    std::printf("\"sampler\": {\"sqrt_spp\": %d}, ", p.sqrt_spp);
    std::printf("\"integrator\": {\"n_incoming\": %u, \"max_bounces\": %d, \"time_sampling\": %d, "
                "\"path_tracing\": 1, \"seed\": %llu, \"scattering_level_all\": 1}, ",
                p.n_incoming, p.max_bounces, int(p.time_sampling), (unsigned long long)p.seed);
    std::printf("\"lights\": [");
    bool first = true;
    for (const auto& l : p.points) {
        std::printf("%s{\"type\": \"point\", ", first ? "" : ", ");
        vec("position", l.position);
        vec("intensity", l.intensities, false);
        std::printf("}");
        first = false;
    }
    for (const auto& l : p.rects) {
        std::printf("%s{\"type\": \"rect\", ", first ? "" : ", ");
        vec("position", l.position);
        vec("target", l.target);
        std::printf("\"size\": [%.9g, %.9g], ", double(l.size[0]), double(l.size[1]));
        vec("intensity", {l.intensity, l.intensity, l.intensity}, false);
        std::printf("}");
        first = false;
    }
    std::printf("], \"medium\": ");
    if (p.has_medium) {
        std::printf("{");
        vec("sigma_a", p.sa);
        vec("sigma_s", p.ss, false);
        std::printf("}");
    } else {
        std::printf("null");
    }
    std::printf(", \"meshes\": [");
    for (size_t k = 0; k < p.meshes.size(); ++k) {
        const MeshArg& m = p.meshes[k];
        float *xyz = nullptr, *nrm = nullptr;
        uint32_t *idx = nullptr, nv = 0, nt = 0;
        if (tr_obj_load(m.file.c_str(), &xyz, &nv, &idx, &nt, &nrm) != TR_OK) {
            std::fprintf(stderr, "%s\n", tr_last_error());
            return 1;
        }
        tr_obj_free(xyz, idx, nrm);
        std::printf("%s{\"file\": \"%s\", \"triangles\": %u, \"two_sided\": %d, \"default_material\": %d, ",
                    k ? ", " : "", m.file.c_str(), nt, int(m.two_sided), int(m.default_material));
        vec("kd", m.kd, false);
        std::printf("}");
    }
    std::printf("]}\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        Params p;
        Args A;
        A.a.assign(argv, argv + argc);
        A.i = 1;
        bool quiet = false;
        for (const auto& a : A.a) quiet = quiet || a == "--print-parsed";
        if (!quiet) {
            for (int i = 0; i < argc; i++) std::fprintf(stdout, "argv[%d]: %s\n", i, argv[i]);
            std::fprintf(stdout, "\n");
        }
        parse(A, p);
        if (p.log_file.empty()) p.log_file = p.image_file + ".log";
        // meshes without a material flag take the grey default, set up after all arguments are read
        // (bunnykiller.cpp:1499-1501), so the last -lambertian-rho applies to all of them
        for (auto& m : p.meshes)
            if (m.default_material) m.kd = {p.grey, p.grey, p.grey};
        // no light on the command line: a unit point light at the camera (bunnykiller.cpp:1636-1645)
        if (p.points.empty() && p.rects.empty()) p.points.emplace_back(p.cam_pos, tr::Vector3{1.f, 1.f, 1.f});
        if (p.print_parsed) return print_parsed(p);

        tr::World w;
        for (const auto& m : p.meshes) w.add_triangle_mesh(m.file, tr::Lambertian(m.kd), m.two_sided);
        for (auto& l : p.points) w.add_light(l);
        for (auto& l : p.rects) w.add_light(l);
        if (p.has_medium) w.set_medium(tr::HomogeneousMedium(p.sa, p.ss));

        FILE* f_log = std::fopen(p.log_file.c_str(), "w+");
        if (f_log) {
            std::fprintf(f_log, "FILM\n====\nName: %s\nSize(x-y-t): %d %d %d\nExposure: %.10f\nt0: %.10f\n",
                         p.image_file.c_str(), p.width, p.height, p.time, double(p.exposure), double(p.offset));
            std::fprintf(f_log, "Sqrt-spp: %d\nSeed: %u\nCamera-unwarp: %s\n\n", p.sqrt_spp, unsigned(p.seed),
                         p.unwarp ? "TRUE" : "FALSE");
        }
        tr::PinholeCamera camera(p.cam_pos, p.cam_focus, p.cam_up, p.view_plane);
        w.freeze();
        tr::StratifiedSampler sampler(p.width, p.height, p.sqrt_spp, true);
        std::unique_ptr<tr::StreakCameraFilm> film;
        if (p.multibounce > 0)
            film.reset(new tr::MultibounceStreakFilm(unsigned(p.width), unsigned(p.height), unsigned(p.time),
                                                     p.exposure, unsigned(p.multibounce), p.unwarp));
        else
            film.reset(new tr::StreakCameraFilm(unsigned(p.width), unsigned(p.height), unsigned(p.time), p.exposure,
                                                p.unwarp));
        film->set_name(p.image_file, "hdr");
        if (p.offset > 0.0f) film->set_offset(p.offset);
        if (p.scanline > -1) {
            sampler.set_scanline(p.scanline);
            film->set_streak(unsigned(p.scanline));
        }
        tr::BidirectionalPathTracing bpt(w, p.n_incoming, f_log, film.get());
        std::printf("Rendering with Bidirectional Path Tracing...\n");
        bpt.set_max_path_size(p.max_bounces);
        bpt.set_path_tracing_only(true);
        // bunnykiller.cpp:1811-1817
        bpt.set_mode(p.time_sampling ? tr::BidirectionalPathTracing::Transient : tr::BidirectionalPathTracing::SteadyState);
        bpt.set_seed(p.seed);
        tr::TransientRenderer engine(f_log, p.time_sampling);
        tr_render_stats st = engine.render(p.image_file.c_str(), w, &bpt, camera, film.get(), &sampler);
        if (f_log) std::fclose(f_log);
        std::printf("{\"paths\": %llu, \"render_ms\": %.3f}\n", (unsigned long long)st.paths, st.render_ms);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
