// Device helpers shared by the look-ahead legalization kernels (lal.hip, the
// multi-kernel path, and lal_batch.hip, the fused one-launch batch): order
// keys of doubles, the bin of a point (simpl_grid.h:120-126) and the record
// of a cell whose bin changed.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace kt {
namespace {

constexpr double kFixScale = 65536.0;  // areas in units of 2^-16

// Unsigned key whose order is the order of the doubles (-0.0 as +0.0).
__device__ __forceinline__ uint64_t orderKey(double v) {
    if (v == 0.0) v = 0.0;  // -0.0 and +0.0 compare equal
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(v));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double fromOrderKey(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double(static_cast<long long>(b));
}

struct GridDev {
    double ox, oy, dx, dy;
    uint32_t nx, ny;
};

// simpl_grid.h:120-126 clampIndex
// (a NaN coordinate maps to bin 0 instead of an undefined conversion)
__device__ __forceinline__ uint32_t clampIndex(double f, uint32_t n) {
    const double hi = static_cast<double>(n) - 1e-9;
    const double c = f >= 0.0 ? (hi < f ? hi : f) : 0.0;
    return static_cast<uint32_t>(c);
}

__device__ __forceinline__ uint32_t binOfPoint(const GridDev &g, double x, double y) {
    return clampIndex((y - g.oy) / g.dy, g.ny) * g.nx + clampIndex((x - g.ox) / g.dx, g.nx);
}

__device__ __forceinline__ uint64_t areaFixOf(const double *__restrict__ area, uint32_t i) {
    return static_cast<uint64_t>(llrint(area[i] * kFixScale));
}

}  // namespace

// Largest batch of the fused one-launch kernel (lal_batch.hip): cells and
// regions; larger batches run on the multi-kernel path.
constexpr uint32_t kLalFusedCells = 1024;
constexpr uint32_t kLalFusedRegions = 64;

// A cell of a batch whose bin changed: the host updates its occupancy, its
// cells per bin and its bin membership lists from these.
struct LalMoved {
    uint32_t oldBin, newBin;
    uint32_t cell, pad;
    unsigned long long area;  // 2^-16 units
};
static_assert(sizeof(LalMoved) == 3 * sizeof(unsigned long long), "a moved record is three words");

}  // namespace kt
