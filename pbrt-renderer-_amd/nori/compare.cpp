// Image metrics of the reference's validation scripts, for comparing renders
// with goldens: validation_tests/compare_exr.py:74-118 (linear-HDR
// statistics, relMSE, tonemapped PSNR) and validation_tests/disney/
// analyze_furnace.py:31-58 (sphere-masked furnace means).
#include "../../include/nori_host.h"

#include <algorithm>
#include <cmath>
#include <limits>

extern "C" {

int nori_image_compare(const float *ours, const float *ref, uint32_t width, uint32_t height, float exposure_ev,
                       nori_image_metrics *out) {
    if (!ours || !ref || !out || width == 0 || height == 0) return NORI_E_ARG;
    const size_t n = (size_t)width * height;
    nori_image_metrics m{};
    double se = 0.0, rel = 0.0, se_tm = 0.0;
    const double scale = std::pow(2.0, (double)exposure_ev);
    auto display = [&](double x) {
        x = std::max(x * scale, 0.0);
        return std::pow(x / (1.0 + x), 1.0 / 2.2);
    };
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const double a = ours[3 * i + c], b = ref[3 * i + c], d = a - b;
            m.mean_ours[c] += a;
            m.mean_ref[c] += b;
            m.mean_abs_diff[c] += std::fabs(d);
            m.max_abs_diff[c] = std::max(m.max_abs_diff[c], std::fabs(d));
            se += d * d;
            rel += d * d / (b * b + 1e-2);
            const double t = display(a) - display(b);
            se_tm += t * t;
        }
    for (int c = 0; c < 3; ++c) {
        m.mean_ours[c] /= (double)n;
        m.mean_ref[c] /= (double)n;
        m.mean_abs_diff[c] /= (double)n;
        m.ratio[c] = m.mean_ref[c] > 1e-10 ? m.mean_ours[c] / m.mean_ref[c] : std::numeric_limits<double>::infinity();
    }
    m.rmse = std::sqrt(se / (3.0 * (double)n));
    m.rel_mse = rel / (3.0 * (double)n);
    const double mse_tm = se_tm / (3.0 * (double)n);
    m.psnr_tonemapped = mse_tm > 1e-10 ? 10.0 * std::log10(1.0 / mse_tm) : std::numeric_limits<double>::infinity();
    *out = m;
    return NORI_OK;
}

int nori_furnace_stats(const float *rgb, uint32_t width, uint32_t height, float fov_deg, float cam_z, float radius,
                       double mean[3], double *lum_mean, uint32_t *count) {
    if (!rgb || !mean || !lum_mean || !count || width == 0 || height == 0) return NORI_E_ARG;
    const double fy = std::tan(fov_deg * M_PI / 180.0 / 2.0), fx = fy * (double)width / (double)height;
    double acc[3] = {0.0, 0.0, 0.0}, lum = 0.0;
    uint32_t k = 0;
    for (uint32_t j = 0; j < height; ++j)
        for (uint32_t i = 0; i < width; ++i) {
            const double u = (i + 0.5) / width * 2.0 - 1.0, v = 1.0 - (j + 0.5) / height * 2.0;
            double dx = u * fx, dy = v * fy, dz = -1.0;
            const double l = std::sqrt(dx * dx + dy * dy + dz * dz);
            dz /= l;
            // |o + t d|^2 = R^2 with o = (0, 0, cam_z): b = o.d, c = |o|^2 - R^2, hit iff b^2 - c > 0
            const double b = cam_z * dz, c = (double)cam_z * cam_z - (double)radius * radius;
            if (b * b - c <= 0.0) continue;
            const float *p = rgb + 3 * ((size_t)j * width + i);
            for (int q = 0; q < 3; ++q) acc[q] += p[q];
            lum += 0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2];
            ++k;
        }
    for (int q = 0; q < 3; ++q) mean[q] = k ? acc[q] / k : 0.0;
    *lum_mean = k ? lum / k : 0.0;
    *count = k;
    return NORI_OK;
}

} // extern "C"
