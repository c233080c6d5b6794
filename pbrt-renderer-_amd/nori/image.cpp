#include "image.h"

#include "common.h"

#include <cstring>
#include <fstream>

namespace nori {

namespace {

ImageRGB readHDR(const std::string &fn) {
    std::ifstream is(fn, std::ios::binary);
    if (!is) throw NoriException("Unable to open \"%s\"", fn);
    std::string line;
    std::getline(is, line);
    if (line.rfind("#?", 0) != 0) throw NoriException("\"%s\": not a Radiance HDR file", fn);
    bool rgbe = false;
    while (std::getline(is, line) && !line.empty())
        if (line.find("FORMAT=32-bit_rle_rgbe") != std::string::npos) rgbe = true;
    if (!rgbe) throw NoriException("\"%s\": unsupported HDR format (expected 32-bit_rle_rgbe)", fn);
    std::getline(is, line);
    char ya[3] = {0}, xa[3] = {0};
    int h = 0, w = 0;
    if (std::sscanf(line.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::string(ya) != "-Y" ||
        std::string(xa) != "+X" || w <= 0 || h <= 0)
        throw NoriException("\"%s\": unsupported HDR resolution line \"%s\"", fn, line);
    ImageRGB img;
    img.width = (uint32_t)w;
    img.height = (uint32_t)h;
    img.rgb.resize((size_t)w * h * 3);
    std::vector<uint8_t> scan((size_t)w * 4);
    auto byte = [&]() {
        int c = is.get();
        if (c == EOF) throw NoriException("\"%s\": truncated pixel data", fn);
        return (uint8_t)c;
    };
    for (int y = 0; y < h; ++y) {
        uint8_t b0 = byte(), b1 = byte(), b2 = byte(), b3 = byte();
        if (w >= 8 && w < 32768 && b0 == 2 && b1 == 2 && ((b2 << 8) | b3) == w && !(b2 & 0x80)) {
            // new-style run-length encoding, one channel after the other
            for (int c = 0; c < 4; ++c) {
                int x = 0;
                while (x < w) {
                    uint8_t n = byte();
                    if (n > 128) {
                        n -= 128;
                        uint8_t v = byte();
                        if (x + n > w) throw NoriException("\"%s\": bad run length", fn);
                        for (int k = 0; k < n; ++k) scan[(size_t)(x++) * 4 + c] = v;
                    } else {
                        if (n == 0 || x + n > w) throw NoriException("\"%s\": bad run length", fn);
                        for (int k = 0; k < n; ++k) scan[(size_t)(x++) * 4 + c] = byte();
                    }
                }
            }
        } else {
            // flat pixels
            scan[0] = b0; scan[1] = b1; scan[2] = b2; scan[3] = b3;
            for (int x = 1; x < w; ++x)
                for (int c = 0; c < 4; ++c) scan[(size_t)x * 4 + c] = byte();
        }
        for (int x = 0; x < w; ++x) {
            const uint8_t *p = &scan[(size_t)x * 4];
            float f = p[3] ? std::ldexp(1.0f, (int)p[3] - (128 + 8)) : 0.0f;
            float *o = &img.rgb[((size_t)y * w + x) * 3];
            o[0] = p[0] * f;
            o[1] = p[1] * f;
            o[2] = p[2] * f;
        }
    }
    return img;
}

ImageRGB readPFM(const std::string &fn) {
    std::ifstream is(fn, std::ios::binary);
    if (!is) throw NoriException("Unable to open \"%s\"", fn);
    std::string magic;
    int w = 0, h = 0;
    float scale = 0.0f;
    is >> magic >> w >> h >> scale;
    is.get();
    if ((magic != "PF" && magic != "Pf") || w <= 0 || h <= 0 || scale == 0.0f)
        throw NoriException("\"%s\": not a PFM file", fn);
    const int ch = magic == "PF" ? 3 : 1;
    std::vector<float> raw((size_t)w * h * ch);
    is.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)(raw.size() * 4));
    if (!is) throw NoriException("\"%s\": truncated pixel data", fn);
    if (scale > 0.0f) // big-endian file
        for (auto &f : raw) {
            uint32_t u;
            std::memcpy(&u, &f, 4);
            u = __builtin_bswap32(u);
            std::memcpy(&f, &u, 4);
        }
    ImageRGB img;
    img.width = (uint32_t)w;
    img.height = (uint32_t)h;
    img.rgb.resize((size_t)w * h * 3);
    for (int y = 0; y < h; ++y) // PFM rows go bottom to top
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c)
                img.rgb[((size_t)y * w + x) * 3 + c] = raw[((size_t)(h - 1 - y) * w + x) * ch + (ch == 3 ? c : 0)];
    return img;
}

void writePFM(const std::string &fn, const ImageRGB &img) {
    std::ofstream os(fn, std::ios::binary);
    if (!os) throw NoriException("Unable to write \"%s\"", fn);
    os << "PF\n" << img.width << " " << img.height << "\n-1\n";
    for (uint32_t y = img.height; y-- > 0;)
        os.write(reinterpret_cast<const char *>(&img.rgb[(size_t)y * img.width * 3]), (std::streamsize)img.width * 12);
    if (!os) throw NoriException("Unable to write \"%s\"", fn);
}

// Single-part scanline OpenEXR, NO_COMPRESSION, channels B, G, R as FLOAT.
void writeEXR(const std::string &fn, const ImageRGB &img) {
    std::string h;
    auto u32 = [&](uint32_t v) { h.append(reinterpret_cast<const char *>(&v), 4); };
    auto i32 = [&](int32_t v) { h.append(reinterpret_cast<const char *>(&v), 4); };
    auto f32 = [&](float v) { h.append(reinterpret_cast<const char *>(&v), 4); };
    auto attr = [&](const char *name, const char *type, uint32_t size) {
        h += name;
        h += '\0';
        h += type;
        h += '\0';
        u32(size);
    };
    u32(20000630u); // magic
    u32(2u);        // version 2, scanline
    attr("channels", "chlist", 3 * (2 + 16) + 1);
    for (const char *c : {"B", "G", "R"}) {
        h += c;
        h += '\0';
        i32(2); // FLOAT
        h.append(4, '\0'); // pLinear + reserved
        i32(1);
        i32(1);
    }
    h += '\0';
    attr("compression", "compression", 1);
    h += '\0';
    attr("dataWindow", "box2i", 16);
    i32(0); i32(0); i32((int32_t)img.width - 1); i32((int32_t)img.height - 1);
    attr("displayWindow", "box2i", 16);
    i32(0); i32(0); i32((int32_t)img.width - 1); i32((int32_t)img.height - 1);
    attr("lineOrder", "lineOrder", 1);
    h += '\0';
    attr("pixelAspectRatio", "float", 4);
    f32(1.0f);
    attr("screenWindowCenter", "v2f", 8);
    f32(0.0f); f32(0.0f);
    attr("screenWindowWidth", "float", 4);
    f32(1.0f);
    h += '\0'; // end of header
    const uint64_t lineBytes = 8 + (uint64_t)img.width * 12;
    uint64_t off = h.size() + 8ull * img.height;
    std::ofstream os(fn, std::ios::binary);
    if (!os) throw NoriException("Unable to write \"%s\"", fn);
    os.write(h.data(), (std::streamsize)h.size());
    for (uint32_t y = 0; y < img.height; ++y) {
        uint64_t o = off + y * lineBytes;
        os.write(reinterpret_cast<const char *>(&o), 8);
    }
    std::vector<float> line((size_t)img.width * 3);
    for (uint32_t y = 0; y < img.height; ++y) {
        int32_t yy = (int32_t)y, sz = (int32_t)(img.width * 12);
        os.write(reinterpret_cast<const char *>(&yy), 4);
        os.write(reinterpret_cast<const char *>(&sz), 4);
        for (int c = 0; c < 3; ++c) // B, G, R planes
            for (uint32_t x = 0; x < img.width; ++x) line[(size_t)c * img.width + x] = img.rgb[((size_t)y * img.width + x) * 3 + (2 - c)];
        os.write(reinterpret_cast<const char *>(line.data()), (std::streamsize)line.size() * 4);
    }
    if (!os) throw NoriException("Unable to write \"%s\"", fn);
}

} // namespace

ImageRGB readImage(const std::string &fn) {
    std::string l = toLower(fn);
    if (endsWith(l, ".hdr") || endsWith(l, ".pic")) return readHDR(fn);
    if (endsWith(l, ".pfm")) return readPFM(fn);
    throw NoriException("\"%s\": unsupported image format (use .hdr or .pfm)", fn);
}

void writeImage(const std::string &fn, const ImageRGB &img) {
    std::string l = toLower(fn);
    if (endsWith(l, ".pfm")) return writePFM(fn, img);
    if (endsWith(l, ".exr")) return writeEXR(fn, img);
    throw NoriException("\"%s\": unsupported image format (use .pfm or .exr)", fn);
}

} // namespace nori
