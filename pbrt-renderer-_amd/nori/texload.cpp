// 8-bit RGBA texture files for the material texture properties of the XML
// scene format (albedoTexture, normalTexture, ...; the reference decodes them
// with stb_image into RGBA8, src/gpu/DXRApp_Textures.cpp:34-45):
//   * PNG (ISO/IEC 15948): bit depths 1-16 of grey, RGB, palette, grey+alpha
//     and RGBA, non-interlaced, zlib streams inflated by the system zlib, the
//     five scanline filters, tRNS transparency for palette images; 16-bit
//     samples keep their high byte;
//   * binary PNM: P5 (grey) and P6 (RGB), maxval up to 255;
//   * baseline JPEG (jpeg.cpp).
// Other formats are not decoded: the caller reports the texture as not loaded,
// as the reference does for a file stb_image refuses.
#include "image.h"

#include <zlib.h>

#include <cstring>
#include <fstream>
#include <iterator>

namespace nori {

namespace {

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

bool fail(std::string &err, const std::string &m) {
    err = m;
    return false;
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return (uint8_t)(pa <= pb && pa <= pc ? a : pb <= pc ? b : c);
}

bool decode_png(const std::vector<uint8_t> &f, uint32_t &w, uint32_t &h, std::vector<uint8_t> &rgba,
                std::string &err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return fail(err, "not a PNG file");
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, pal, trns;
    bool ihdr = false;
    for (size_t p = 8; p + 12 <= f.size();) {
        const uint32_t len = be32(&f[p]);
        if (p + 12 + (size_t)len > f.size()) return fail(err, "truncated PNG chunk");
        const char *type = (const char *)&f[p + 4];
        const uint8_t *d = &f[p + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) return fail(err, "bad IHDR");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
            ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            pal.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        }
        p += 12 + (size_t)len;
    }
    if (!ihdr || w == 0 || h == 0 || w > (1u << 15) || h > (1u << 15)) return fail(err, "bad PNG size");
    if (interlace != 0) return fail(err, "interlaced PNG is not supported");
    int channels;
    switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: return fail(err, "bad PNG colour type");
    }
    const bool depth_ok = ctype == 0 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)
                          : ctype == 3 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8)
                                       : (depth == 8 || depth == 16);
    if (!depth_ok) return fail(err, "bad PNG bit depth");
    if (ctype == 3 && pal.size() < 3) return fail(err, "palette PNG without PLTE");
    const size_t bits_px = (size_t)channels * depth, stride = ((size_t)w * bits_px + 7) / 8;
    const size_t bpp = std::max<size_t>(1, bits_px / 8);
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf out_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK || out_len != raw.size())
        return fail(err, "corrupt PNG image data");
    // unfilter in place, row by row
    std::vector<uint8_t> img(stride * h);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t *src = &raw[y * (stride + 1) + 1];
        uint8_t *cur = &img[y * stride];
        const uint8_t *prev = y ? &img[(y - 1) * stride] : nullptr;
        for (size_t x = 0; x < stride; ++x) {
            const int a = x >= bpp ? cur[x - bpp] : 0, b = prev ? prev[x] : 0, c = prev && x >= bpp ? prev[x - bpp] : 0;
            int v = src[x];
            switch (ft) {
            case 0: break;
            case 1: v += a; break;
            case 2: v += b; break;
            case 3: v += (a + b) / 2; break;
            case 4: v += paeth(a, b, c); break;
            default: return fail(err, "bad PNG filter type");
            }
            cur[x] = (uint8_t)v;
        }
    }
    auto sample = [&](const uint8_t *row, uint32_t x, int ch) -> uint32_t {
        if (depth == 8) return row[(size_t)x * channels + ch];
        if (depth == 16) return row[2 * ((size_t)x * channels + ch)]; // high byte
        const size_t bit = (size_t)x * depth; // sub-byte depths: one channel
        return (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1u);
    };
    rgba.assign((size_t)w * h * 4, 255);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t *row = &img[y * stride];
        for (uint32_t x = 0; x < w; ++x) {
            uint8_t *o = &rgba[4 * ((size_t)y * w + x)];
            if (ctype == 3) {
                const uint32_t k = sample(row, x, 0);
                if (3 * k + 2 >= pal.size()) return fail(err, "PNG palette index out of range");
                o[0] = pal[3 * k];
                o[1] = pal[3 * k + 1];
                o[2] = pal[3 * k + 2];
                o[3] = k < trns.size() ? trns[k] : 255;
                continue;
            }
            uint32_t v[4];
            for (int ch = 0; ch < channels; ++ch) {
                v[ch] = sample(row, x, ch);
                if (depth < 8) v[ch] = v[ch] * 255u / ((1u << depth) - 1u);
            }
            if (channels <= 2) {
                o[0] = o[1] = o[2] = (uint8_t)v[0];
                if (channels == 2) o[3] = (uint8_t)v[1];
            } else {
                o[0] = (uint8_t)v[0];
                o[1] = (uint8_t)v[1];
                o[2] = (uint8_t)v[2];
                if (channels == 4) o[3] = (uint8_t)v[3];
            }
        }
    }
    return true;
}

bool decode_pnm(const std::vector<uint8_t> &f, uint32_t &w, uint32_t &h, std::vector<uint8_t> &rgba,
                std::string &err) {
    if (f.size() < 3 || f[0] != 'P' || (f[1] != '5' && f[1] != '6')) return fail(err, "not a binary PNM file");
    const int channels = f[1] == '6' ? 3 : 1;
    size_t p = 2;
    uint32_t vals[3];
    for (int k = 0; k < 3; ++k) {
        for (;;) { // whitespace and comments
            while (p < f.size() && std::isspace(f[p])) ++p;
            if (p < f.size() && f[p] == '#')
                while (p < f.size() && f[p] != '\n') ++p;
            else
                break;
        }
        uint64_t v = 0;
        if (p >= f.size() || !std::isdigit(f[p])) return fail(err, "bad PNM header");
        while (p < f.size() && std::isdigit(f[p]) && v < (1u << 20)) v = v * 10 + (f[p++] - '0');
        vals[k] = (uint32_t)v;
    }
    ++p; // one whitespace byte before the raster
    w = vals[0];
    h = vals[1];
    const uint32_t maxval = vals[2];
    if (w == 0 || h == 0 || w > (1u << 15) || h > (1u << 15) || maxval == 0 || maxval > 255)
        return fail(err, "unsupported PNM size or maxval");
    if (p + (size_t)w * h * channels > f.size()) return fail(err, "truncated PNM raster");
    rgba.assign((size_t)w * h * 4, 255);
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c)
            rgba[4 * i + c] = (uint8_t)(f[p + i * channels + (channels == 3 ? c : 0)] * 255u / maxval);
    return true;
}

} // namespace

bool loadTextureRGBA8(const std::string &path, uint32_t &width, uint32_t &height, std::vector<uint8_t> &rgba,
                      std::string &err) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(err, "cannot open file");
    std::vector<uint8_t> f((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G')
        return decode_png(f, width, height, rgba, err);
    if (f.size() >= 2 && f[0] == 'P' && (f[1] == '5' || f[1] == '6')) return decode_pnm(f, width, height, rgba, err);
    if (f.size() >= 2 && f[0] == 0xFF && f[1] == 0xD8) return decodeJpegRGBA8(f, width, height, rgba, err);
    return fail(err, "unsupported image format (PNG, baseline JPEG and binary PNM are decoded)");
}

} // namespace nori
