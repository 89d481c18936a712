// Image output of the streak films: Radiance RGBE (.hdr) and raw FP32 dumps, the two formats the
// reference writes through Imaging::save (src/Image/ImageIO.h:147-194).
//
// RGBE follows Greg Ward's Radiance picture format ("Real Pixels", Graphics Gems II, 1991): a
// text header, a resolution line "-Y <height> +X <width>", then 4-byte pixels (mantissas r, g, b
// scaled by a shared exponent byte e: value = m * 2^(e - 136)). The files carry the same bytes as
// the reference's: its writer (the vendored stb_image_write, stb_image_write.h:588-640, 687-705)
// emits the header lines "#?RADIANCE", "FORMAT=32-bit_rle_rgbe", "EXPOSURE=" with a fixed value,
// and every scanline as flat pixels (its run-length branch is disabled), so a reader of the
// reference's output reads these files alike.
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

int write_rgbe(const char* path, const float* rgb, int width, int height)
{
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(TR_ERR_INVALID, std::string("Error: failed to open output file: \"") + path + "\"");
    char res[128];
    const int n = std::snprintf(res, sizeof res, "-Y %d +X %d\n", height, width);
    const std::string head = std::string("#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=          1.0000000000000\n\n") +
                             std::string(res, size_t(n));
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    std::vector<uint8_t> px(size_t(width) * 4);
    for (int y = 0; y < height && ok; ++y) {
        const float* row = rgb + size_t(y) * size_t(width) * 3;
        for (int x = 0; x < width; ++x) rgbe_from_float(row[3 * x], row[3 * x + 1], row[3 * x + 2], &px[4 * size_t(x)]);
        ok = std::fwrite(px.data(), 1, px.size(), f) == px.size();
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
