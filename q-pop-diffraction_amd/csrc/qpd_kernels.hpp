// The three pass kernels of one diffraction volume (gfx950).
//
// The reference forms, per atom a, four fields fE_a, fE_a u_x, fE_a u_y, fE_a u_z,
// transforms each with a 3-D FFT and adds phi_a(k) [FE - i k . FU] to the amplitude
// (src/python/diffraction.py:97-109, src/heffte/kernels.cuh:55-111).  phi_a(k) and
// the voxel shape factor S(k) factor into one term per axis, and -i k_d acts on one
// axis only, so both commute with the transforms along the other axes.  A volume is
// therefore three passes, one per memory dim, each "FFT along the dim, multiply by a
// per-index coefficient, sum into fewer outputs":
//
//   pass 0 (contiguous dim):  inputs computed from P, U and the weights (the source);
//                             per atom [E, U_d0, U_d1, U_d2] -> per class of atoms with
//                             equal (R_d1, R_d2): [A, B_d1, B_d2]
//   pass 1 (middle dim):      per class [A, B_d1, B_d2] -> per class of equal R_d2: [C, D]
//   pass 2 (outer dim):       per class [C, D] -> amplitude -> fftshifted max(|.|^2, 1e-50)
//
// In every pass a chunk of M+1 inputs feeds the group's M outputs:
//   out_0 += c(k) F[in_0] + c(k) (-i k) F[in_1],  out_j += c(k) F[in_{j+1}] (j >= 1),
// with c(k) a per-chunk coefficient table (exp(-i k R) s(k), and exp(-i G.R_a)/(V n)
// in pass 0).  A plain transform is the case M+1 = 1 with c = 1.
//
// Offsets inside one field are 32-bit (the host rejects grids of 2^31 cells or more);
// field base pointers are block-uniform 64-bit values.
#pragma once

#include "qpd_device.hpp"

namespace qpd {

// per (atom, phase): Re f, Im f, su[d][k] (d, k = 0..2), g[k] = sum_d qC_d su[d][k]
constexpr int kAtomTab = 14;

template <typename T>
struct PassArgs {
    int N;                    // transform length along this pass
    int nst;                  // its number of stages
    const Stage* stages;      // its stage list (device memory)
    const cx<T>* W;           // twiddles exp(-2 pi i m / N), N entries
    const cx<T>* coef;        // coefficient tables, [table][N]
    const T* k;               // wavenumber of each index along this dim, nm^-1
    const int* chunk_src;     // source of each chunk: atom (pass 0) or input class
    const int* chunk_coef;    // coefficient table of each chunk, -1 = identity
    const int* group_first;   // chunks of group g: [group_first[g], group_first[g+1])
    int ngroups;
    int nin;                  // inputs per chunk (M + 1, or 1 for a plain transform)
    int nout;                 // outputs per group (M)
    int nf;                   // inputs resident in LDS at once
    uint32_t n0, n1, n2;      // memory extents, n0 fastest
    uint32_t ncell;
    int nl;                   // rows (contiguous pass) or TX (strided pass) per block
    FastDiv divN;             // division by N (element -> row)
    FastDiv divL1, divL2, divL3, divL4;  // division by nfc * nl, nfc = 1..4
    // strided pass: lines along dim `dim` (1 or 2)
    int dim;
    int log2TX;
    uint32_t ntiles0;
    uint32_t S;               // element stride along the pass
    uint32_t OS;              // stride of the outer index
    // split transforms (fast strided kernels): a length N = N1 N2 runs as N2-point
    // sub-transforms over rows q + N1 y2 (q < N1), then N1-point ones over rows N1 q + y1
    // (q < N2) written to rows q + N2 k1.  The outer index of a block is q + nq * outer;
    // element idx of a line is read from row q in_qmul + idx in_step and written to row
    // q out_qmul + idx out_step; its full-length frequency index is q c_qmul + idx c_step.
    uint32_t nq;
    FastDiv divQ;
    uint32_t in_qmul, in_step, out_qmul, out_step, c_qmul, c_step;
    int Nc;                   // length of the k and coefficient tables (the full length)
    const cx<T>* Wsplit;      // exp(-2 pi i m / Nc), m < Nc: twiddles between the two steps
    int nfields;              // fields transformed in place by the first step
    // pass 1 over source groups (fast_strided_mix): nsrc groups of 3 inputs; for group s
    // and z-class z the y-classes pair_y[pair_first[s NZ + z] .. pair_first[s NZ + z + 1])
    // contribute with column factor mixc[(y nsrc + s) n0 + i0]
    int nsrc;
    int ncoef;                // coefficient tables of the pass (the y-classes), staged in LDS by fast_mix_held
    const int* pair_first;
    const int* pair_y;
    const cx<T>* mixc;
    // data
    const cx<T>* in;
    cx<T>* out;
    double* I;
    // source (pass 0 only)
    const T* P;
    const double* U;
    const signed char* phase;
    const T* fraction;
    const T* region;
    const T* atomtab;         // [atom][phase][kAtomTab]
    double qC0, qC1, qC2;
    int n_phase;
    int n_atom_tab;           // entries of atomtab (n_atom * n_phase * kAtomTab)
    int ax0, ax1, ax2;        // physical axis of memory dims 0, 1, 2
};

// Workgroups are dispatched to the 8 XCDs round robin (block b to XCD b % 8), and each
// XCD has its own L2.  Strided passes read rows of G columns; neighbouring column tiles
// share 128-byte lines, so the logical block b / 8 + (b % 8) * (nblocks / 8) puts runs of
// consecutive tiles on one XCD, where the second read of a line hits L2.  The last
// nblocks % 8 blocks keep their index (a bijection on [0, nblocks)).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
    constexpr uint32_t kXcd = 8;
    const uint32_t per = nblocks / kXcd;
    return b < per * kXcd ? (b % kXcd) * per + b / kXcd : b;
}

// v[i] for a run-time i in 0..2 without indexing a private array.
__device__ __forceinline__ double pick3(double v0, double v1, double v2, int i) {
    return i == 0 ? v0 : (i == 1 ? v1 : v2);
}

template <typename T>
__device__ __forceinline__ FastDiv lines_div(const PassArgs<T>& a, int nfc) {
    return nfc == 1 ? a.divL1 : (nfc == 2 ? a.divL2 : (nfc == 3 ? a.divL3 : a.divL4));
}

// ---------------------------------------------------------------------------
// Source of one cell for one atom (kernels.cuh:55-86; diffraction.py:100-104):
//   E = sum_p w_p f_pa e^{-i qC.u_ap},  U_j = sum_p w_p f_pa e^{-i qC.u_ap} u_ap[ax_j],
//   u_ap = U + su[:, :, p, a] . P.
// The phase qC.u_ap is formed in FP64; see phase_sincos.
template <typename T>
struct CellInputs {
    T P0, P1, P2;
    double U0, U1, U2;
    T rg;
    int ph;
};

template <typename T, bool IO32 = false>
__device__ __forceinline__ void cell_load(const PassArgs<T>& a, uint32_t off, bool valid, CellInputs<T>& c) {
    using TR = io_t<T, IO32>;
    c.ph = -1;
    c.rg = (T)1;
    c.P0 = c.P1 = c.P2 = (T)0;
    c.U0 = c.U1 = c.U2 = 0.0;
    if (!valid) return;
    const TR* P = reinterpret_cast<const TR*>(a.P);
    c.ph = a.phase[off];
    if (a.region) c.rg = (T)reinterpret_cast<const TR*>(a.region)[off];
    c.P0 = (T)P[off];
    c.P1 = (T)P[a.ncell + off];
    c.P2 = (T)P[(size_t)2 * a.ncell + off];
    c.U0 = a.U[off];
    c.U1 = a.U[a.ncell + off];
    c.U2 = a.U[(size_t)2 * a.ncell + off];
}

template <typename T, bool IO32 = false>
__device__ __forceinline__ void cell_source(const PassArgs<T>& a, const T* atomtab, const CellInputs<T>& ci,
                                            uint32_t off, int atom, cx<T>& E, cx<T>& V0, cx<T>& V1, cx<T>& V2) {
    E = V0 = V1 = V2 = {(T)0, (T)0};
    if (ci.ph < 0) return;
    const int p_lo = a.fraction ? 0 : ci.ph;
    const int p_hi = a.fraction ? a.n_phase : ci.ph + 1;
    for (int p = p_lo; p < p_hi; ++p) {
        T w = ci.rg;
        if (a.fraction) w *= (T)reinterpret_cast<const io_t<T, IO32>*>(a.fraction)[(size_t)p * a.ncell + off];
        if (w == (T)0) continue;
        const T* tab = atomtab + (atom * a.n_phase + p) * kAtomTab;
        const double ux = ci.U0 + (double)(tab[2] * ci.P0 + tab[3] * ci.P1 + tab[4] * ci.P2);
        const double uy = ci.U1 + (double)(tab[5] * ci.P0 + tab[6] * ci.P1 + tab[7] * ci.P2);
        const double uz = ci.U2 + (double)(tab[8] * ci.P0 + tab[9] * ci.P1 + tab[10] * ci.P2);
        const double t = a.qC0 * ux + a.qC1 * uy + a.qC2 * uz;
        T sn, cs;
        phase_sincos(t, sn, cs);
        const T fr = tab[0] * w, fi = tab[1] * w;
        // w f e^{-i t} = w (fr + i fi)(cs - i sn)
        const cx<T> e = {fr * cs + fi * sn, fi * cs - fr * sn};
        E = E + e;
        V0 = V0 + cscale(e, (T)pick3(ux, uy, uz, a.ax0));
        V1 = V1 + cscale(e, (T)pick3(ux, uy, uz, a.ax1));
        V2 = V2 + cscale(e, (T)pick3(ux, uy, uz, a.ax2));
    }
}

// The same source for one-hot phases (no fraction field) in factored form: the fields are
// e, e u_d0, e u_d1, e u_d2 with e = w f e^{-i t} and u_d the displacement along memory
// dim d (cell_source forms exactly these products).
template <typename T>
__device__ __forceinline__ void cell_source_eu(const PassArgs<T>& a, const T* atomtab, const CellInputs<T>& ci,
                                               int atom, cx<T>& e, T& u0, T& u1, T& u2) {
    e = {(T)0, (T)0};
    u0 = u1 = u2 = (T)0;
    if (ci.ph < 0) return;
    const T* tab = atomtab + (atom * a.n_phase + ci.ph) * kAtomTab;
    const double ux = ci.U0 + (double)(tab[2] * ci.P0 + tab[3] * ci.P1 + tab[4] * ci.P2);
    const double uy = ci.U1 + (double)(tab[5] * ci.P0 + tab[6] * ci.P1 + tab[7] * ci.P2);
    const double uz = ci.U2 + (double)(tab[8] * ci.P0 + tab[9] * ci.P1 + tab[10] * ci.P2);
    const double t = a.qC0 * ux + a.qC1 * uy + a.qC2 * uz;
    T sn, cs;
    phase_sincos(t, sn, cs);
    const T fr = tab[0] * ci.rg, fi = tab[1] * ci.rg;
    e = {fr * cs + fi * sn, fi * cs - fr * sn};
    u0 = (T)pick3(ux, uy, uz, a.ax0);
    u1 = (T)pick3(ux, uy, uz, a.ax1);
    u2 = (T)pick3(ux, uy, uz, a.ax2);
}

// Input jj of a chunk, already multiplied by the chunk coefficient, into the outputs:
// 0 -> out0, 1 -> out0 times -i k, j >= 2 -> out_{j-1}.
template <typename T, int NOUT>
__device__ __forceinline__ void accumulate_term(int jj, cx<T> t, T kv, cx<T>* acc) {
    if (jj == 0) {
        acc[0] = acc[0] + t;
    } else if (jj == 1) {
        acc[0] = acc[0] + cscale(negi(t), kv);
    } else if constexpr (NOUT >= 2) {
        if (jj == 2) {
            acc[1] = acc[1] + t;
        } else if constexpr (NOUT >= 3) {
            acc[2] = acc[2] + t;
        }
    }
}

template <typename T>
__device__ __forceinline__ cx<T> chunk_coef(const PassArgs<T>& a, int ct, int idx) {
    return ct >= 0 ? a.coef[ct * a.Nc + idx] : cx<T>{(T)1, (T)0};
}

// ---------------------------------------------------------------------------
// Pass along the contiguous dim: block = nl consecutive rows of n0 points.  RS: the set of
// radices of the axis (fft_tile).
// LDS tile: point idx of line (f, row) at tile[idx * (nf*nl) + f*nl + row].
// Element e = tid + m*nthr of a field's tile is (row, idx) = divmod(e, N).
// IO32: FP64 arithmetic on FP32 arrays (inputs, P and weights, outputs).
template <typename T, int EPT, int NOUT, bool SOURCE, int RS = kRadixAll, bool IO32 = false>
__global__ __launch_bounds__(512) void pass_contig(PassArgs<T> a) {
    using TS = io_t<T, IO32>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int N = a.N;
    const int nl = a.nl;
    const uint32_t nrows = a.n1 * a.n2;
    const uint32_t row0 = blockIdx.x * (uint32_t)nl;
    const int tileN = nl * N;
    const int stride = a.nf * nl;
    const uint32_t cell0 = row0 * (uint32_t)N;

    CellInputs<T> ci[SOURCE ? EPT : 1];
    if constexpr (SOURCE) {
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int e = tid + m * nthr;
            const int row = (int)fdiv((uint32_t)e, a.divN);
            cell_load<T, IO32>(a, cell0 + (uint32_t)e, e < tileN && row0 + row < nrows, ci[m]);
        }
    }

    for (int g = 0; g < a.ngroups; ++g) {
        cx<T> acc[EPT][NOUT];
#pragma unroll
        for (int m = 0; m < EPT; ++m)
#pragma unroll
            for (int o = 0; o < NOUT; ++o) acc[m][o] = {(T)0, (T)0};

        for (int c = uload(a.group_first, g); c < uload(a.group_first, g + 1); ++c) {
            const int src = uload(a.chunk_src, c);
            const int ct = uload(a.chunk_coef, c);
            for (int j0 = 0; j0 < a.nin; j0 += a.nf) {
                const int nfc = min(a.nf, a.nin - j0);
                const cx<TS>* inbase =
                    SOURCE ? nullptr : reinterpret_cast<const cx<TS>*>(a.in) + ((size_t)src * a.nin + j0) * a.ncell + cell0;
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const int e = tid + m * nthr;
                    if (e >= tileN) continue;
                    const int row = (int)fdiv((uint32_t)e, a.divN);
                    const int idx = e - row * N;
                    cx<T>* dst = tile + idx * stride + row;
                    if constexpr (SOURCE) {
                        cx<T> v0, v1, v2, v3;
                        cell_source<T, IO32>(a, a.atomtab, ci[m], cell0 + (uint32_t)e, src, v0, v1, v2, v3);
                        for (int f = 0; f < nfc; ++f) {
                            const int jj = j0 + f;
                            dst[f * nl] = jj == 0 ? v0 : (jj == 1 ? v1 : (jj == 2 ? v2 : v3));
                        }
                    } else {
                        const bool valid = row0 + row < nrows;
                        for (int f = 0; f < nfc; ++f)
                            dst[f * nl] = valid ? cx_as<T>(inbase[(size_t)f * a.ncell + e]) : cx<T>{(T)0, (T)0};
                    }
                }
                __syncthreads();
                fft_tile<T, RS, EPT >= 16 ? kMaxRegsLong : kMaxRegs>(tile, a.nst, a.stages, a.W, stride, nfc * nl, lines_div(a, nfc), tid, nthr);
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const int e = tid + m * nthr;
                    if (e >= tileN) continue;
                    const int row = (int)fdiv((uint32_t)e, a.divN);
                    const int idx = e - row * N;
                    const cx<T> cf = chunk_coef(a, ct, idx);
                    const T kv = a.k[idx];
                    const cx<T>* srcp = tile + idx * stride + row;
                    for (int f = 0; f < nfc; ++f) accumulate_term<T, NOUT>(j0 + f, cmul(cf, srcp[f * nl]), kv, acc[m]);
                }
                __syncthreads();
            }
        }
        cx<TS>* obase = reinterpret_cast<cx<TS>*>(a.out) + (size_t)g * NOUT * a.ncell + cell0;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int e = tid + m * nthr;
            if (e >= tileN) continue;
            const int row = (int)fdiv((uint32_t)e, a.divN);
            if (row0 + row >= nrows) continue;
#pragma unroll
            for (int q = 0; q < NOUT; ++q) obase[(size_t)q * a.ncell + e] = cx_as<TS>(acc[m][q]);
        }
    }
}

// ---------------------------------------------------------------------------
// Pass along a strided dim (1 or 2): block = TX consecutive i0 x one outer index.
// LDS tile: point idx of line (f, tx) at tile[idx * (nf*TX) + f*TX + tx].  A thread
// keeps one column tx = tid % TX and the points idx = tid / TX + m * (nthr / TX).
// INTENSITY: the single output is the amplitude; write fftshifted max(|A|^2, 1e-50).
// IO32: FP64 arithmetic on complex FP32 inputs and outputs.
template <typename T, int EPT, int NOUT, bool INTENSITY, int RS = kRadixAll, bool IO32 = false>
__global__ __launch_bounds__(512) void pass_strided(PassArgs<T> a) {
    using TI = io_t<T, IO32>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int N = a.N;
    const int TX = 1 << a.log2TX;
    const int stride = a.nf * TX;
    // narrow tiles share HBM lines with their neighbours: keep neighbours on one XCD
    const uint32_t blk = TX * sizeof(cx<TI>) < 128 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t t0 = blk % a.ntiles0;
    const uint32_t outer = blk / a.ntiles0;
    const int tx = tid & (TX - 1);
    const int r0 = tid >> a.log2TX;
    const int rstep = nthr >> a.log2TX;
    const uint32_t i0 = t0 * (uint32_t)TX + (uint32_t)tx;
    const bool col_ok = i0 < a.n0;
    const uint32_t colbase = i0 + outer * a.OS;  // offset of point idx = 0 of this column

    for (int g = 0; g < a.ngroups; ++g) {
        cx<T> acc[EPT][NOUT];
#pragma unroll
        for (int m = 0; m < EPT; ++m)
#pragma unroll
            for (int o = 0; o < NOUT; ++o) acc[m][o] = {(T)0, (T)0};

        for (int c = uload(a.group_first, g); c < uload(a.group_first, g + 1); ++c) {
            const int src = uload(a.chunk_src, c);
            const int ct = uload(a.chunk_coef, c);
            for (int j0 = 0; j0 < a.nin; j0 += a.nf) {
                const int nfc = min(a.nf, a.nin - j0);
                const cx<TI>* inbase = reinterpret_cast<const cx<TI>*>(a.in) + ((size_t)src * a.nin + j0) * a.ncell;
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const int idx = r0 + m * rstep;
                    if (idx >= N) continue;
                    const uint32_t off = colbase + (uint32_t)idx * a.S;
                    cx<T>* dst = tile + idx * stride + tx;
                    for (int f = 0; f < nfc; ++f)
                        dst[f * TX] = col_ok ? cx_as<T>(inbase[(size_t)f * a.ncell + off]) : cx<T>{(T)0, (T)0};
                }
                __syncthreads();
                fft_tile<T, RS, EPT >= 16 ? kMaxRegsLong : kMaxRegs>(tile, a.nst, a.stages, a.W, stride, nfc * TX, lines_div(a, nfc), tid, nthr);
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const int idx = r0 + m * rstep;
                    if (idx >= N) continue;
                    const cx<T> cf = chunk_coef(a, ct, idx);
                    const T kv = a.k[idx];
                    const cx<T>* srcp = tile + idx * stride + tx;
                    for (int f = 0; f < nfc; ++f) accumulate_term<T, NOUT>(j0 + f, cmul(cf, srcp[f * TX]), kv, acc[m]);
                }
                __syncthreads();
            }
        }
        if (!col_ok) continue;
        if constexpr (INTENSITY) {
            // fftshift: index i of an axis of length n goes to (i + n/2) mod n
            const uint32_t s0 = (i0 + a.n0 / 2) >= a.n0 ? i0 + a.n0 / 2 - a.n0 : i0 + a.n0 / 2;
            const uint32_t nA = a.dim == 2 ? a.n2 : a.n1;  // axis of idx
            const uint32_t nB = a.dim == 2 ? a.n1 : a.n2;  // axis of outer
            const uint32_t sB = (outer + nB / 2) >= nB ? outer + nB / 2 - nB : outer + nB / 2;
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const int idx = r0 + m * rstep;
                if (idx >= N) continue;
                const uint32_t sA = ((uint32_t)idx + nA / 2) >= nA ? idx + nA / 2 - nA : idx + nA / 2;
                const uint32_t s1 = a.dim == 2 ? sB : sA;
                const uint32_t s2 = a.dim == 2 ? sA : sB;
                const double re = (double)acc[m][0].x, im = (double)acc[m][0].y;
                a.I[(size_t)s0 + (size_t)a.n0 * ((size_t)s1 + (size_t)a.n1 * s2)] = fmax(re * re + im * im, 1e-50);
            }
        } else {
            cx<TI>* obase = reinterpret_cast<cx<TI>*>(a.out) + (size_t)g * NOUT * a.ncell;
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const int idx = r0 + m * rstep;
                if (idx >= N) continue;
                const uint32_t off = colbase + (uint32_t)idx * a.S;
#pragma unroll
                for (int q = 0; q < NOUT; ++q) obase[(size_t)q * a.ncell + off] = cx_as<TI>(acc[m][q]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pass 1 over source groups for any length and precision (the run-time-radix form of
// fast_strided_mix): the three inputs [A_s, B_s, C_s] of every source group s are
// transformed in the tile, and every z-class z receives
//   out_{z,0} = sum_s sum_{Y in z} c_Y(k1) X_{Y,s}(k0) [F1 A_s - i k1 F1 B_s],
//   out_{z,1} = sum_s sum_{Y in z} c_Y(k1) X_{Y,s}(k0) F1 C_s,
// the outputs of the class form's pass 1.
// IO32: FP64 arithmetic on complex FP32 inputs and outputs.
template <typename T, int EPT, int NZ, int RS = kRadixAll, bool IO32 = false>
__global__ __launch_bounds__(512) void pass_strided_mix(PassArgs<T> a) {
    using TI = io_t<T, IO32>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int N = a.N;
    const int TX = 1 << a.log2TX;
    const int stride = a.nf * TX;
    // narrow tiles share HBM lines with their neighbours: keep neighbours on one XCD
    const uint32_t blk = TX * sizeof(cx<TI>) < 128 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t t0 = blk % a.ntiles0;
    const uint32_t outer = blk / a.ntiles0;
    const int tx = tid & (TX - 1);
    const int r0 = tid >> a.log2TX;
    const int rstep = nthr >> a.log2TX;
    const uint32_t i0 = t0 * (uint32_t)TX + (uint32_t)tx;
    const bool col_ok = i0 < a.n0;
    const uint32_t colbase = i0 + outer * a.OS;

    cx<T> acc[EPT][NZ][2];
#pragma unroll
    for (int m = 0; m < EPT; ++m)
#pragma unroll
        for (int z = 0; z < NZ; ++z) acc[m][z][0] = acc[m][z][1] = {(T)0, (T)0};
    for (int s = 0; s < a.nsrc; ++s) {
        for (int j0 = 0; j0 < 3; j0 += a.nf) {
            const int nfc = min(a.nf, 3 - j0);
            const cx<TI>* inbase = reinterpret_cast<const cx<TI>*>(a.in) + ((size_t)s * 3 + j0) * a.ncell;
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const int idx = r0 + m * rstep;
                if (idx >= N) continue;
                const uint32_t off = colbase + (uint32_t)idx * a.S;
                cx<T>* dst = tile + idx * stride + tx;
                for (int f = 0; f < nfc; ++f)
                    dst[f * TX] = col_ok ? cx_as<T>(inbase[(size_t)f * a.ncell + off]) : cx<T>{(T)0, (T)0};
            }
            __syncthreads();
            fft_tile<T, RS, EPT >= 16 ? kMaxRegsLong : kMaxRegs>(tile, a.nst, a.stages, a.W, stride, nfc * TX,
                                                                  lines_div(a, nfc), tid, nthr);
            // each transformed field goes into the sums as it is read from the tile (no
            // per-group values are held across fields): F A and -i k F B into out_{z,0},
            // F C into out_{z,1}
            if (col_ok) {
                for (int f = 0; f < nfc; ++f) {
                    const int jj = j0 + f;
#pragma unroll
                    for (int z = 0; z < NZ; ++z) {
                        for (int p = uload(a.pair_first, s * NZ + z); p < uload(a.pair_first, s * NZ + z + 1); ++p) {
                            const int y = uload(a.pair_y, p);
                            const cx<T> xc = a.mixc[(size_t)(y * a.nsrc + s) * a.n0 + i0];
#pragma unroll
                            for (int m = 0; m < EPT; ++m) {
                                const int idx = r0 + m * rstep;
                                if (idx >= N) continue;
                                cx<T> t = cmul(cmul(a.coef[(size_t)y * a.Nc + idx], xc), tile[idx * stride + f * TX + tx]);
                                if (jj == 1) t = cscale(negi(t), a.k[idx]);
                                if (jj <= 1)
                                    acc[m][z][0] = acc[m][z][0] + t;
                                else
                                    acc[m][z][1] = acc[m][z][1] + t;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    if (!col_ok) return;
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int idx = r0 + m * rstep;
        if (idx >= N) continue;
        const uint32_t off = colbase + (uint32_t)idx * a.S;
#pragma unroll
        for (int z = 0; z < NZ; ++z)
#pragma unroll
            for (int oq = 0; oq < 2; ++oq)
                reinterpret_cast<cx<TI>*>(a.out)[(size_t)(2 * z + oq) * a.ncell + off] = cx_as<TI>(acc[m][z][oq]);
    }
}

}  // namespace qpd
