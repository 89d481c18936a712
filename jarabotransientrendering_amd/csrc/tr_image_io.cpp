// Image output of the streak films: Radiance RGBE (.hdr) and raw FP32 dumps, the two formats the
// reference writes through Imaging::save (src/Image/ImageIO.h:147-194).
//
// RGBE follows Greg Ward's Radiance picture format ("Real Pixels", Graphics Gems II, 1991): a
// text header, a resolution line "-Y <height> +X <width>", then scanlines of 4-byte pixels
// (mantissas r, g, b scaled by a shared exponent byte e: value = m * 2^(e - 136)). Scanlines of
// 8..32767 pixels use the run-length encoding of the format's "new" scheme: a 4-byte marker
// (2, 2, width >> 8, width & 255), then each of the four byte planes as runs (count > 128:
// count - 128 repeats of one byte) and literal spans (count <= 128 bytes).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "tr_api.h"
#include "tr_internal.h"

namespace tri {

void rgbe_from_float(float r, float g, float b, uint8_t out[4])
{
    float m = r > g ? r : g;
    m = m > b ? m : b;
    if (!(m >= 1e-32f)) {  // also zero for NaN
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e = 0;
    float scale = float(std::frexp(m, &e) * 256.0 / m);
    // mantissas truncate toward zero; a negative channel (never produced by a film) wraps
    // modulo 256 as the reference's x86 build does
    out[0] = uint8_t(int32_t(r * scale));
    out[1] = uint8_t(int32_t(g * scale));
    out[2] = uint8_t(int32_t(b * scale));
    out[3] = uint8_t(e + 128);
}

namespace {

// one byte plane of a scanline in the run-length scheme
void encode_plane(const uint8_t* v, int n, std::vector<uint8_t>& out)
{
    int x = 0;
    while (x < n) {
        // find the next run of at least 4 equal bytes starting at or after x
        int run_start = x, run_len = 0;
        while (run_start < n) {
            run_len = 1;
            while (run_start + run_len < n && run_len < 127 && v[run_start + run_len] == v[run_start]) ++run_len;
            if (run_len >= 4) break;
            run_start += run_len;
            run_len = 0;
        }
        // literal bytes before the run, in spans of at most 128
        while (x < run_start) {
            int k = run_start - x;
            if (k > 128) k = 128;
            out.push_back(uint8_t(k));
            out.insert(out.end(), v + x, v + x + k);
            x += k;
        }
        if (run_len >= 4) {
            out.push_back(uint8_t(128 + run_len));
            out.push_back(v[run_start]);
            x = run_start + run_len;
        }
    }
}

}  // namespace

int write_rgbe(const char* path, const float* rgb, int width, int height)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(TR_ERR_INVALID, std::string("Error: failed to open output file: \"") + path + "\"");
    std::string head = "#?RADIANCE\n# transient path tracer\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(height) +
                       " +X " + std::to_string(width) + "\n";
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    std::vector<uint8_t> px(size_t(width) * 4), planes(size_t(width) * 4), out;
    for (int y = 0; y < height && ok; ++y) {
        const float* row = rgb + size_t(y) * size_t(width) * 3;
        for (int x = 0; x < width; ++x) rgbe_from_float(row[3 * x], row[3 * x + 1], row[3 * x + 2], &px[4 * size_t(x)]);
        out.clear();
        if (width < 8 || width > 32767) {
            out.assign(px.begin(), px.end());
        } else {
            out.push_back(2);
            out.push_back(2);
            out.push_back(uint8_t(width >> 8));
            out.push_back(uint8_t(width & 255));
            for (int c = 0; c < 4; ++c) {
                for (int x = 0; x < width; ++x) planes[size_t(c) * width + x] = px[4 * size_t(x) + c];
                encode_plane(&planes[size_t(c) * width], width, out);
            }
        }
        ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(TR_ERR_INVALID, std::string("Error: failed to write to output file: ") + path);
    return TR_OK;
}

int write_raw(const char* path, const float* data, size_t count)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(TR_ERR_INVALID, std::string("Error: failed to open output file: \"") + path + "\"");
    bool ok = std::fwrite(data, sizeof(float), count, f) == count;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(TR_ERR_INVALID, std::string("Error: failed to write to output file: ") + path);
    return TR_OK;
}

}  // namespace tri

extern "C" int tr_image_save(const char* path, const float* rgb, int width, int height)
{
    if (!path || !rgb || width <= 0 || height <= 0) return tri::fail(TR_ERR_INVALID, "tr_image_save: bad argument");
    std::string p(path);
    size_t dot = p.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : p.substr(dot + 1);
    if (ext == "hdr") return tri::write_rgbe(path, rgb, width, height);
    if (ext == "raw") return tri::write_raw(path, rgb, size_t(width) * size_t(height) * 3);
    return tri::fail(TR_ERR_INVALID, "Error: unsupported output file format: \"" + ext + "\"");
}
