// render_scene: renders a scene description through the C++ host classes (tr_bunnykiller.hpp),
// wired the way the reference's main() wires its objects (bunnykiller.cpp:1504-1827).
//
// Input: a text scene file, one directive per line (the same format the procedural scene
// module writes):
//   film W H T exposure offset scanline unwarp
//   camera px py pz fx fy fz ux uy uz fov_degrees
//   render sqrt_spp n_incoming max_bounces seed
//   material r g b
//   tri x0 y0 z0 x1 y1 z1 x2 y2 z2 material_index
//   point x y z r g b
//   rect cx cy cz tx ty tz size_x size_y intensity
//   medium sigma_a(rgb) sigma_s(rgb)
//   multibounce nb_bounces
//   timesampling            (the integrator's Transient sampling mode, -time-sampling)
//   out name_prefix
// Output: <prefix>_%04d.raw streak slices, <prefix>.log, and one JSON line of render statistics.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "tr_bunnykiller.hpp"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s scene.txt\n", argv[0]);
        return 2;
    }
    try {
        std::ifstream in(argv[1]);
        if (!in) throw std::runtime_error(std::string("cannot open ") + argv[1]);
        int W = 64, H = 64, T = 256, scanline = -1, unwarp = 0, sqrt_spp = 2, n_inc = 1, max_b = 4;
        float exposure = 0.01f, offset = 0.f;
        double cam[10] = {0.5, 0.5, -1.2, 0.5, 0.5, 0.5, 0, 1, 0, 45};
        unsigned long long seed = 0;
        std::vector<tr::Vector3> mats;
        std::vector<std::vector<float>> mesh_xyz;  // one mesh per run of equal material
        std::vector<int> mesh_mat;
        std::vector<tr::PointLightSource> points;
        std::vector<tr::RectangularAreaLightSource> rects;
        bool has_medium = false;
        tr::Vector3 sa{}, ss{};
        int multibounce = 0;
        bool time_sampling = false;
        std::string out = "film";
        std::string line;
        // consecutive triangles of one material form one mesh, as a -name-mesh per material
        int cur_mat = -1;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string key;
            if (!(ls >> key) || key[0] == '#') continue;
            if (key == "film") {
                ls >> W >> H >> T >> exposure >> offset >> scanline >> unwarp;
            } else if (key == "camera") {
                for (double& v : cam) ls >> v;
            } else if (key == "render") {
                ls >> sqrt_spp >> n_inc >> max_b >> seed;
            } else if (key == "material") {
                tr::Vector3 m;
                ls >> m[0] >> m[1] >> m[2];
                mats.push_back(m);
            } else if (key == "tri") {
                float v[9];
                int m;
                for (float& x : v) ls >> x;
                ls >> m;
                if (m < 0 || size_t(m) >= mats.size()) throw std::runtime_error("tri: bad material index");
                if (m != cur_mat) {
                    mesh_xyz.emplace_back();
                    mesh_mat.push_back(m);
                    cur_mat = m;
                }
                mesh_xyz.back().insert(mesh_xyz.back().end(), v, v + 9);
            } else if (key == "point") {
                float p[6];
                for (float& x : p) ls >> x;
                points.emplace_back(tr::Vector3{p[0], p[1], p[2]}, tr::Vector3{p[3], p[4], p[5]});
            } else if (key == "rect") {
                float r[9];
                for (float& x : r) ls >> x;
                rects.emplace_back(tr::Vector3{r[0], r[1], r[2]}, tr::Vector3{r[3], r[4], r[5]},
                                   std::array<float, 2>{r[6], r[7]}, r[8]);
            } else if (key == "medium") {
                ls >> sa[0] >> sa[1] >> sa[2] >> ss[0] >> ss[1] >> ss[2];
                has_medium = true;
            } else if (key == "multibounce") {
                ls >> multibounce;
            } else if (key == "timesampling") {
                time_sampling = true;
            } else if (key == "out") {
                ls >> out;
            } else {
                throw std::runtime_error("unknown directive " + key);
            }
        }

        tr::World w;
        for (size_t k = 0; k < mesh_xyz.size(); ++k) {
            std::vector<uint32_t> idx(mesh_xyz[k].size() / 3);
            for (size_t i = 0; i < idx.size(); ++i) idx[i] = uint32_t(i);
            w.add_triangle_mesh(mesh_xyz[k], idx, tr::Lambertian(mats[size_t(mesh_mat[k])]));
        }
        for (auto& p : points) w.add_light(p);
        for (auto& r : rects) w.add_light(r);
        if (has_medium) w.set_medium(tr::HomogeneousMedium(sa, ss));

        FILE* f_log = std::fopen((out + ".log").c_str(), "w+");
        tr::PinholeCamera camera = tr::PinholeCamera::from_fov(
            {float(cam[0]), float(cam[1]), float(cam[2])}, {float(cam[3]), float(cam[4]), float(cam[5])},
            {float(cam[6]), float(cam[7]), float(cam[8])}, float(cam[9]));
        w.freeze();
        tr::StratifiedSampler sampler(W, H, sqrt_spp, true);
        std::unique_ptr<tr::StreakCameraFilm> film_ptr;
        if (multibounce > 0)
            film_ptr.reset(new tr::MultibounceStreakFilm(unsigned(W), unsigned(H), unsigned(T), exposure,
                                                         unsigned(multibounce), unwarp != 0));
        else
            film_ptr.reset(new tr::StreakCameraFilm(unsigned(W), unsigned(H), unsigned(T), exposure, unwarp != 0));
        tr::StreakCameraFilm& film = *film_ptr;
        film.set_name(out, "raw");
        if (offset > 0.0f) film.set_offset(offset);
        if (scanline > -1) {
            sampler.set_scanline(scanline);
            film.set_streak(unsigned(scanline));
        }
        tr::BidirectionalPathTracing bpt(w, unsigned(n_inc), f_log, &film);
        bpt.set_max_path_size(max_b);
        bpt.set_path_tracing_only(true);
        bpt.set_mode(time_sampling ? tr::BidirectionalPathTracing::Transient : tr::BidirectionalPathTracing::SteadyState);
        bpt.set_seed(seed);
        tr::TransientRenderer engine(f_log, time_sampling);
        tr_render_stats st = engine.render(out.c_str(), w, &bpt, camera, &film, &sampler);
        if (f_log) std::fclose(f_log);
        std::printf("{\"paths\": %llu, \"extension_rays\": %llu, \"shadow_rays\": %llu, \"contributions\": %llu, "
                    "\"render_ms\": %.3f}\n",
                    (unsigned long long)st.paths, (unsigned long long)st.extension_rays,
                    (unsigned long long)st.shadow_rays, (unsigned long long)st.contributions, st.render_ms);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
