// Axis reversal of a grid's fields: the memory order (a0 fastest, a1, a2) becomes (a2 fastest,
// a1, a0).  The CLI lays FerroDyn fields out x fastest; a volume whose x rows are longer than
// every compile-time source kernel runs faster z fastest (the source pass then transforms the
// short z rows), so its inputs are reversed before the plan and its intensity after.
//
// For each component and each a1 the (a2, a0) plane is transposed through a 32 x 32 LDS tile
// (one column of padding against bank conflicts): reads run along a0 and writes along a2, both
// 32 elements a row.  Blocks of 32 x 8 threads; a 1-D grid of tiles, so no grid dimension limit
// binds the a1 and component extents.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qpd {

constexpr int kTrTile = 32;
constexpr int kTrRows = 8;

template <typename E>
__global__ __launch_bounds__(kTrTile* kTrRows) void reverse_axes_kernel(const E* __restrict__ src, E* __restrict__ dst,
                                                                         uint32_t n0, uint32_t n1, uint32_t n2,
                                                                         uint32_t t0, uint32_t t2) {
    __shared__ E tile[kTrTile][kTrTile + 1];
    // block -> (tile along a0, tile along a2, a1, component)
    uint64_t b = blockIdx.x;
    const uint32_t i0 = (uint32_t)(b % t0);
    b /= t0;
    const uint32_t i2 = (uint32_t)(b % t2);
    b /= t2;
    const uint32_t a1 = (uint32_t)(b % n1);
    const uint64_t comp = b / n1;
    const size_t plane = (size_t)n0 * n1 * n2;
    const E* s = src + comp * plane;
    E* d = dst + comp * plane;
    const uint32_t tx = threadIdx.x, ty = threadIdx.y;
    const uint32_t x0 = i0 * kTrTile, z0 = i2 * kTrTile;
    // read: element (a0 = x0 + tx, a2 = z0 + r) at a0 + n0 (a1 + n1 a2)
    for (uint32_t r = ty; r < kTrTile; r += kTrRows) {
        const uint32_t a0 = x0 + tx, a2 = z0 + r;
        if (a0 < n0 && a2 < n2) tile[r][tx] = s[(size_t)a0 + (size_t)n0 * ((size_t)a1 + (size_t)n1 * a2)];
    }
    __syncthreads();
    // write: element (a2 = z0 + tx, a0 = x0 + r) at a2 + n2 (a1 + n1 a0)
    for (uint32_t r = ty; r < kTrTile; r += kTrRows) {
        const uint32_t a2 = z0 + tx, a0 = x0 + r;
        if (a0 < n0 && a2 < n2) d[(size_t)a2 + (size_t)n2 * ((size_t)a1 + (size_t)n1 * a0)] = tile[tx][r];
    }
}

}  // namespace qpd
