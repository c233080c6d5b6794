#include "volume.h"

#include <algorithm>
#include <cmath>
#include <thread>

namespace ng {

namespace {
// Voxels a trilinear lookup reads for normalised coordinates in [k/m, (k+1)/m]
// of an axis of n voxels: floor(u n - 0.5) .. floor(u n - 0.5) + 1, clamped.
void voxel_range(uint32_t k, uint32_t m, uint32_t n, int &lo, int &hi) {
    double a = (double)k / m * n - 0.5, b = (double)(k + 1) / m * n - 0.5;
    lo = std::max(0, (int)std::floor(a));
    hi = std::min((int)n - 1, (int)std::floor(b) + 1);
}
} // namespace

void build_majorant(uint32_t W, uint32_t H, uint32_t D, const float *density, uint32_t brick, uint32_t &mW,
                    uint32_t &mH, uint32_t &mD, std::vector<float> &out) {
    const uint32_t B = brick ? brick : 16;
    mW = (W + B - 1) / B;
    mH = (H + B - 1) / B;
    mD = (D + B - 1) / B;
    out.assign((size_t)mW * mH * mD, 0.0f);
    auto slab = [&](uint32_t z_begin, uint32_t z_end) {
        for (uint32_t bz = z_begin; bz < z_end; ++bz)
            for (uint32_t by = 0; by < mH; ++by)
                for (uint32_t bx = 0; bx < mW; ++bx) {
                    int x0, x1, y0, y1, z0, z1;
                    voxel_range(bx, mW, W, x0, x1);
                    voxel_range(by, mH, H, y0, y1);
                    voxel_range(bz, mD, D, z0, z1);
                    float m = 0.0f;
                    for (int z = z0; z <= z1; ++z)
                        for (int y = y0; y <= y1; ++y) {
                            const float *row = density + (size_t)W * ((size_t)y + (size_t)H * z);
                            for (int x = x0; x <= x1; ++x) m = std::max(m, row[x]);
                        }
                    out[(size_t)bx + (size_t)mW * ((size_t)by + (size_t)mH * bz)] = m;
                }
    };
    const uint32_t nth = std::max(1u, std::min(mD, std::min(16u, std::thread::hardware_concurrency())));
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < nth; ++t) th.emplace_back(slab, mD * t / nth, mD * (t + 1) / nth);
    for (auto &x : th) x.join();
}

} // namespace ng
