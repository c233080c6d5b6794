// nori_compare: the reference's compare_exr.py / analyze_furnace.py reports
// for two renders (.pfm or .hdr), through libnori_host.so.
//   nori_compare ours ref [--exposure EV]
//   nori_compare --furnace image [ref]   (unit sphere at the origin, camera at
//                                         z = 4, vertical fov 45 degrees)
#include "../../include/nori_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace {

struct Image {
    float *rgb = nullptr;
    uint32_t w = 0, h = 0;
    ~Image() { nori_image_free(rgb); }
};

bool load(const char *path, Image &img) {
    char err[512];
    if (nori_image_read(path, &img.rgb, &img.w, &img.h, err, sizeof err) != NORI_OK) {
        std::fprintf(stderr, "nori_compare: %s\n", err);
        return false;
    }
    return true;
}

int furnace(int argc, char **argv) {
    Image a, b;
    if (argc < 3 || !load(argv[2], a)) return 2;
    double m[3], lum;
    uint32_t n;
    nori_furnace_stats(a.rgb, a.w, a.h, 45.0f, 4.0f, 1.0f, m, &lum, &n);
    std::printf("Image          : %u x %u    sphere pixels = %u / %u\n", a.w, a.h, n, a.w * a.h);
    std::printf("sphere mean R,G,B = %.6f %.6f %.6f   luminance mean = %.6f\n", m[0], m[1], m[2], lum);
    if (argc > 3) {
        if (!load(argv[3], b)) return 2;
        double r[3], rl;
        uint32_t rn;
        nori_furnace_stats(b.rgb, b.w, b.h, 45.0f, 4.0f, 1.0f, r, &rl, &rn);
        double dev = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double q = m[c] / (r[c] > 1e-10 ? r[c] : 1e-10);
            dev = std::max(dev, std::abs(q - 1.0));
            std::printf("ratio %c = %.4f\n", "RGB"[c], q);
        }
        const double ql = lum / (rl > 1e-10 ? rl : 1e-10);
        dev = std::max(dev, std::abs(ql - 1.0));
        std::printf("luminance ratio = %.4f\n%s max deviation %.2f%%\n", ql,
                    dev < 0.02 ? "PASS" : dev < 0.05 ? "CLOSE" : "FAIL", 100.0 * dev);
    }
    return 0;
}

} // namespace

int main(int argc, char **argv) {
    if (argc >= 2 && std::strcmp(argv[1], "--furnace") == 0) return furnace(argc, argv);
    if (argc < 3) {
        std::fprintf(stderr, "usage: nori_compare ours ref [--exposure EV] | nori_compare --furnace image [ref]\n");
        return 2;
    }
    float ev = 0.0f;
    if (argc >= 5 && std::strcmp(argv[3], "--exposure") == 0) ev = (float)std::atof(argv[4]);
    Image a, b;
    if (!load(argv[1], a) || !load(argv[2], b)) return 2;
    if (a.w != b.w || a.h != b.h) {
        std::fprintf(stderr, "nori_compare: size mismatch %ux%u vs %ux%u\n", a.w, a.h, b.w, b.h);
        return 1;
    }
    nori_image_metrics m;
    nori_image_compare(a.rgb, b.rgb, a.w, a.h, ev, &m);
    std::printf("Resolution      : %u x %u\n", a.w, a.h);
    std::printf("%8s  %10s  %10s  %8s  %10s  %10s\n", "channel", "ours_mean", "ref_mean", "ratio", "mean|diff|",
                "max|diff|");
    for (int c = 0; c < 3; ++c)
        std::printf("%8c  %10.4f  %10.4f  %8.4f  %10.4f  %10.4f\n", "RGB"[c], m.mean_ours[c], m.mean_ref[c], m.ratio[c],
                    m.mean_abs_diff[c], m.max_abs_diff[c]);
    std::printf("Linear RMSE     : %.5f\nrelMSE          : %.5f\nTonemap PSNR    : %.2f dB\n", m.rmse, m.rel_mse,
                m.psnr_tonemapped);
    return 0;
}
