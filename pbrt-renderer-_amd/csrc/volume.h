// Host-side preparation of participating media for the device.
#pragma once
#include <cstdint>
#include <vector>

namespace ng {

// Per-brick majorant grid of a W x H x D density grid (x fastest): bricks of
// `brick` voxels, mW = ceil(W / brick) etc.; each entry is the maximum density
// over every voxel a trilinear lookup inside the brick's world region can read
// (the brick's voxels plus a one-voxel apron), so density / majorant <= 1.
void build_majorant(uint32_t W, uint32_t H, uint32_t D, const float *density, uint32_t brick, uint32_t &mW,
                    uint32_t &mH, uint32_t &mD, std::vector<float> &out);

} // namespace ng
