// Sums of a 3-D float64 array over each of its memory dims (gfx950), for reciprocal-space
// maps: on the rectilinear q grid a `sum` map is the intensity summed over the third axis.
// Every sum runs in a fixed order, so results do not depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qpd {

// S2[i0 + d0 i1] = sum over i2 of A[i0 + d0 (i1 + d1 i2)]: one thread per (i0, i1); lanes take
// consecutive i0, so each step is one coalesced row.
__global__ __launch_bounds__(256) void project_outer(const double* __restrict__ A, double* __restrict__ S, uint32_t d0,
                                                    uint32_t d1, uint32_t d2) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= d0 * d1) return;
    const size_t plane = (size_t)d0 * d1;
    double s = 0.0;
    for (uint32_t i2 = 0; i2 < d2; ++i2) s += A[t + plane * i2];
    S[t] = s;
}

// S1[i0 + d0 i2] = sum over i1 of A[i0 + d0 (i1 + d1 i2)]: one thread per (i0, i2).
__global__ __launch_bounds__(256) void project_middle(const double* __restrict__ A, double* __restrict__ S, uint32_t d0,
                                                     uint32_t d1, uint32_t d2) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= d0 * d2) return;
    const uint32_t i0 = t % d0, i2 = t / d0;
    const double* row = A + (size_t)i0 + (size_t)d0 * d1 * i2;
    double s = 0.0;
    for (uint32_t i1 = 0; i1 < d1; ++i1) s += row[(size_t)d0 * i1];
    S[t] = s;
}

// S0[r] = sum over i0 of row r (r = i1 + d1 i2): one 64-lane wave per row, four rows per
// block; each lane sums a strided part in order, then a fixed xor-shuffle tree.
__global__ __launch_bounds__(256) void project_inner(const double* __restrict__ A, double* __restrict__ S, uint32_t d0,
                                                    uint32_t rows) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t r = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (r >= rows) return;  // whole waves leave together
    const double* row = A + (size_t)r * d0;
    double s = 0.0;
    for (uint32_t i = lane; i < d0; i += 64u) s += row[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) S[r] = s;
}

}  // namespace qpd
