// Field ingest (gfx950): raw FerroDyn fields -> plan inputs, with the u_z reference ramp.
//
// One thread per (x, y) column of the crop walks z over the full stack: it sums dc of the
// column's materials from z = 0 (the reference ramp is cumulative over the whole stack,
// cli.py:207-208) and, inside the crop, converts P and u and writes P, U and the phase
// index.  Lanes take consecutive x, so every z step is one coalesced row per wave; the
// loads of several z steps are in flight together, and the per-material dc and phase
// index are read from LDS.
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

constexpr int kIngestMaterials = 256;  // materials whose dc and phase index are staged in LDS
constexpr int kIngestZ = 4;            // z steps whose loads are issued together

template <typename TP, typename TU, typename TO>
__global__ __launch_bounds__(256) void ingest_columns(IngestArgs a, TO* __restrict__ P_out, double* __restrict__ U_out,
                                                      signed char* __restrict__ phase_out) {
    // each operation rounds on its own, as NumPy's do: no fused multiply-adds
#pragma clang fp contract(off)
    __shared__ double dc_l[kIngestMaterials];
    __shared__ signed char ph_l[kIngestMaterials];
    const bool staged = a.n_material <= kIngestMaterials;
    if (staged) {
        for (int m = threadIdx.x; m < a.n_material; m += 256) {
            dc_l[m] = a.dc[m];
            ph_l[m] = a.phase_of[m];
        }
    }
    __syncthreads();
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= a.nc) return;
    const TP* __restrict__ P0 = static_cast<const TP*>(a.P0);
    const TP* __restrict__ P1 = static_cast<const TP*>(a.P1);
    const TP* __restrict__ P2 = static_cast<const TP*>(a.P2);
    const TU* __restrict__ U0 = static_cast<const TU*>(a.U0);
    const TU* __restrict__ U1 = static_cast<const TU*>(a.U1);
    const TU* __restrict__ U2 = static_cast<const TU*>(a.U2);
    const int32_t* __restrict__ mat = a.mat;
    const bool has_u = U0 != nullptr;
    const TO pdiv = (TO)a.P_div;
    const int64_t z_end = a.z0 + a.nz;
    double sum = 0.0;
    // kIngestZ z steps at a time: their loads are issued before any of their arithmetic;
    // the dc sum still runs in z order
    for (int64_t z = 0; z < a.nz_full; z += kIngestZ) {
        int m[kIngestZ];
        TP p[kIngestZ][3];
        TU u[kIngestZ][3];
#pragma unroll
        for (int q = 0; q < kIngestZ; ++q) {
            const int64_t zz = z + q;
            m[q] = zz < a.nz_full ? mat[(size_t)c + (size_t)a.nc * (size_t)zz] : -1;
            p[q][0] = p[q][1] = p[q][2] = (TP)0;
            u[q][0] = u[q][1] = u[q][2] = (TU)0;
            if (zz >= a.z0 && zz < z_end) {
                const size_t i = (size_t)c + (size_t)a.nc * (size_t)(zz - a.z0);
                p[q][0] = P0[i];
                p[q][1] = P1[i];
                p[q][2] = P2[i];
                if (has_u) {
                    u[q][0] = U0[i];
                    u[q][1] = U1[i];
                    u[q][2] = U2[i];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kIngestZ; ++q) {
            const int64_t zz = z + q;
            if (zz >= a.nz_full) break;
            const int mq = m[q];
            const bool known = (unsigned)mq < (unsigned)a.n_material;
            const double d = known ? (staged ? dc_l[mq] : a.dc[mq]) : 0.0;
            sum += d;
            if (zz < a.z0 || zz >= z_end) continue;
            const size_t i = (size_t)c + (size_t)a.nc * (size_t)(zz - a.z0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                TO pk = (TO)p[q][k];
                if (a.P_div != 1.0) pk = pk / pdiv;
                P_out[(size_t)k * a.n + i] = pk;
                double uk = has_u ? (double)u[q][k] / a.U_div : 0.0;
                if (k == 2) uk += (sum - 0.5 * d) * a.dz;
                U_out[(size_t)k * a.n + i] = uk;
            }
            phase_out[i] = known ? (staged ? ph_l[mq] : a.phase_of[mq]) : (signed char)-1;
        }
    }
}

}  // namespace qpd
