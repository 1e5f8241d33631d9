// hdr.cpp — Radiance RGBE (.hdr) decoding into RGBA32F (hvx_hdr_decode of
// include/hvx.h): the input format of ResourcesIBL's environment
// (VulkanImage::CreateFromHDR, HelloVulkan/Source/Vulkan/VulkanImage.cpp:120-146,
// which loads with stb_image's stbi_loadf after
// stbi_set_flip_vertically_on_load(true); stb_image is not vendored in the
// reference, its behaviour is restated from the Radiance file format).
//
// Format: a text header starting "#?RADIANCE" or "#?RGBE", lines up to an
// empty line, which must declare FORMAT=32-bit_rle_rgbe; then a resolution
// line "-Y <height> +X <width>" (rows top to bottom, columns left to right);
// then the scanlines. A scanline of width 8..32767 whose first four bytes are
// 2, 2, width >> 8, width & 255 is run-length encoded: the four channel
// planes in turn, as runs (a count byte c > 128 followed by one byte repeated
// c - 128 times) and literals (c <= 128 followed by c bytes). Any other
// scanline, and the rest of the image after one, is flat RGBE quadruples.
// A texel (r, g, b, e) decodes to (r, g, b) * 2^(e - 136), or 0 when e == 0;
// alpha is 1. With flip_vertically the last file row becomes row 0.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hvx.h"

namespace {

struct Reader {
    const uint8_t* p;
    size_t n, i = 0;
    bool line(std::string& out) {
        out.clear();
        while (i < n) {
            const char c = (char)p[i++];
            if (c == '\n') return true;
            out.push_back(c);
            if (out.size() > 4096) return false;
        }
        return false;
    }
};

void put(float* dst, const uint8_t q[4]) {
    if (q[3] == 0) {
        dst[0] = dst[1] = dst[2] = 0.0f;
    } else {
        const float f = std::ldexp(1.0f, (int)q[3] - 136);
        dst[0] = (float)q[0] * f;
        dst[1] = (float)q[1] * f;
        dst[2] = (float)q[2] * f;
    }
    dst[3] = 1.0f;
}

}  // namespace

extern "C" int hvx_hdr_decode(const uint8_t* data, size_t size, int flip_vertically, uint32_t* width,
                              uint32_t* height, float* rgba) {
    if (!data || !width || !height) return HVX_ERR_NULL;
    Reader r{data, size};
    std::string s;
    if (!r.line(s) || (s.rfind("#?RADIANCE", 0) != 0 && s.rfind("#?RGBE", 0) != 0)) return HVX_ERR_INVALID_ARG;
    bool format = false;
    for (;;) {
        if (!r.line(s)) return HVX_ERR_INVALID_ARG;
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") format = true;
    }
    if (!format) return HVX_ERR_UNSUPPORTED;
    if (!r.line(s)) return HVX_ERR_INVALID_ARG;
    int h = 0, w = 0;
    char tail = 0;
    if (std::sscanf(s.c_str(), "-Y %d +X %d%c", &h, &w, &tail) != 2) return HVX_ERR_UNSUPPORTED;
    if (h <= 0 || w <= 0 || h > (1 << 16) || w > (1 << 16)) return HVX_ERR_INVALID_ARG;
    *width = (uint32_t)w;
    *height = (uint32_t)h;
    if (!rgba) return HVX_OK;
    auto row_ptr = [&](int y) { return rgba + (size_t)(flip_vertically ? h - 1 - y : y) * w * 4; };
    std::vector<uint8_t> line((size_t)w * 4);
    bool flat = w < 8 || w >= 32768;
    for (int y = 0; y < h; ++y) {
        float* dst = row_ptr(y);
        if (!flat) {
            if (r.i + 4 > r.n) return HVX_ERR_INVALID_ARG;
            const uint8_t* c = r.p + r.i;
            if (c[0] != 2 || c[1] != 2 || (c[2] & 0x80)) {
                flat = true;  // the rest of the image is flat
            } else {
                if (((int)c[2] << 8 | c[3]) != w) return HVX_ERR_INVALID_ARG;
                r.i += 4;
                for (int ch = 0; ch < 4; ++ch) {
                    int x = 0;
                    while (x < w) {
                        if (r.i >= r.n) return HVX_ERR_INVALID_ARG;
                        int cnt = r.p[r.i++];
                        if (cnt > 128) {
                            cnt -= 128;
                            if (cnt > w - x || r.i >= r.n) return HVX_ERR_INVALID_ARG;
                            const uint8_t v = r.p[r.i++];
                            for (int k = 0; k < cnt; ++k) line[(size_t)(x++) * 4 + ch] = v;
                        } else {
                            if (cnt == 0 || cnt > w - x || r.i + cnt > r.n) return HVX_ERR_INVALID_ARG;
                            for (int k = 0; k < cnt; ++k) line[(size_t)(x++) * 4 + ch] = r.p[r.i++];
                        }
                    }
                }
                for (int x = 0; x < w; ++x) put(dst + (size_t)x * 4, &line[(size_t)x * 4]);
                continue;
            }
        }
        if (r.i + (size_t)w * 4 > r.n) return HVX_ERR_INVALID_ARG;
        for (int x = 0; x < w; ++x) put(dst + (size_t)x * 4, r.p + r.i + (size_t)x * 4);
        r.i += (size_t)w * 4;
    }
    return HVX_OK;
}
