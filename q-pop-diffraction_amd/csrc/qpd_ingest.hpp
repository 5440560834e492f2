// Field ingest (gfx950): raw FerroDyn fields -> plan inputs, with the u_z reference ramp.
//
// One thread per (x, y) column of the crop walks z over the full stack: it sums dc of the
// column's materials from z = 0 (the reference ramp is cumulative over the whole stack,
// cli.py:207-208) and, inside the crop, converts P and u and writes P, U and the phase
// index.  Lanes take consecutive x, so every z step is one coalesced row per wave.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qpd {

struct IngestArgs {
    uint32_t nc;        // columns of the crop (nx ny)
    uint32_t n;         // cells of the crop
    int64_t nz_full, z0, nz;
    const void* P0;
    const void* P1;
    const void* P2;
    const void* U0;
    const void* U1;
    const void* U2;
    double P_div, U_div, dz;
    const int32_t* mat;
    int n_material;
    const double* dc;
    const signed char* phase_of;
};

template <typename TP, typename TU, typename TO>
__global__ __launch_bounds__(256) void ingest_columns(IngestArgs a, TO* __restrict__ P_out, double* __restrict__ U_out,
                                                      signed char* __restrict__ phase_out) {
    // each operation rounds on its own, as NumPy's do: no fused multiply-adds
#pragma clang fp contract(off)
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= a.nc) return;
    const TP* P[3] = {static_cast<const TP*>(a.P0), static_cast<const TP*>(a.P1), static_cast<const TP*>(a.P2)};
    const TU* U[3] = {static_cast<const TU*>(a.U0), static_cast<const TU*>(a.U1), static_cast<const TU*>(a.U2)};
    const bool has_u = a.U0 != nullptr;
    const TO pdiv = (TO)a.P_div;
    double sum = 0.0;
    for (int64_t z = 0; z < a.nz_full; ++z) {
        const int m = a.mat[(size_t)c + (size_t)a.nc * (size_t)z];
        const bool known = (unsigned)m < (unsigned)a.n_material;
        const double d = known ? a.dc[m] : 0.0;
        sum += d;
        if (z < a.z0 || z >= a.z0 + a.nz) continue;
        const size_t i = (size_t)c + (size_t)a.nc * (size_t)(z - a.z0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            TO p = (TO)P[k][i];
            if (a.P_div != 1.0) p = p / pdiv;
            P_out[(size_t)k * a.n + i] = p;
            double u = has_u ? (double)U[k][i] / a.U_div : 0.0;
            if (k == 2) u += (sum - 0.5 * d) * a.dz;
            U_out[(size_t)k * a.n + i] = u;
        }
        phase_out[i] = known ? a.phase_of[m] : (signed char)-1;
    }
}

}  // namespace qpd
