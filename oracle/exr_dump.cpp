// Test infrastructure: decodes an OpenEXR image (the reference's Mitsuba
// goldens are 32-bit float, PIZ-compressed) with the tinyexr header the
// reference vendors (src/gpu/tinyexr.h, compiled from where it lies; see
// ref.mk) and writes  int32 width, int32 height, float32 RGB[height][width]
// to the output file.  Used only to turn goldens into test fixtures.
#include <zlib.h>
#define TINYEXR_USE_MINIZ 0
#define TINYEXR_USE_STB_ZLIB 0
#define TINYEXR_IMPLEMENTATION
#include "tinyexr.h"

#include <cstdint>
#include <cstdio>
#include <vector>

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: exr_dump in.exr out.bin\n");
        return 2;
    }
    float *rgba = nullptr;
    int w = 0, h = 0;
    const char *err = nullptr;
    if (LoadEXR(&rgba, &w, &h, argv[1], &err) != TINYEXR_SUCCESS) {
        std::fprintf(stderr, "%s: %s\n", argv[1], err ? err : "load failed");
        if (err) FreeEXRErrorMessage(err);
        return 1;
    }
    std::vector<float> rgb((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c) rgb[3 * i + c] = rgba[4 * i + c];
    std::free(rgba);
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    const int32_t dims[2] = {w, h};
    const bool ok = std::fwrite(dims, 4, 2, f) == 2 && std::fwrite(rgb.data(), 4, rgb.size(), f) == rgb.size();
    return std::fclose(f) == 0 && ok ? 0 : 1;
}
