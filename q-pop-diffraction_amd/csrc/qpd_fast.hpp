// Register-staged Stockham passes with compile-time transform lengths (gfx950).
//
// Same three passes and the same chunk / group / coefficient structure as
// qpd_kernels.hpp, for lengths whose radix sequence is known at compile time.
// A line of N points is handled by TPL = max_s(N / R_s) threads; thread j owns
// butterfly j of every stage.  The first stage takes its inputs x[j + r N/R_0]
// straight from global memory (or from the source computation) into registers,
// the last stage leaves its outputs X[j + r N/R_last] in registers, where they are
// multiplied by the chunk coefficient and accumulated; only the hand-offs between
// stages go through LDS.  Each thread transforms the F inputs of a chunk at once,
// so its twiddles are loaded once per stage and reused F times.
//
// LDS layouts (8-byte complex points, four-byte banks):
//   contiguous pass: line-major, point idx of line l at l * PITCH + idx + idx / 8
//                    (the padding breaks the stride-R pattern of stage outputs);
//   strided pass:    point-major, point idx of line l in row idx (+ idx / 8 when the first
//                    radix is a power of two) of L (+ 1 otherwise) points, with the lanes
//                    of a wave walking consecutive columns (lines); see Sizes::PAD8.
#pragma once

#include <type_traits>

#include "qpd_kernels.hpp"

namespace qpd {

template <int N_, int R0_, int R1_, int R2_ = 1, int R3_ = 1>
struct Sizes {
    static constexpr int N = N_;
    static constexpr int NST = R3_ > 1 ? 4 : (R2_ > 1 ? 3 : (R1_ > 1 ? 2 : 1));
    static constexpr int R[4] = {R0_, R1_, R2_, R3_};
    static constexpr int rad(int s) { return s == 0 ? R0_ : s == 1 ? R1_ : s == 2 ? R2_ : R3_; }
    static constexpr int nr(int s) { return N_ / rad(s); }
    static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * rad(s - 1); }
    static constexpr int cmax(int a, int b) { return a > b ? a : b; }
    static constexpr int TPL = cmax(cmax(nr(0), NST > 1 ? nr(1) : 0), cmax(NST > 2 ? nr(2) : 0, NST > 3 ? nr(3) : 0));
    static constexpr int RMAX = cmax(cmax(R0_, R1_), cmax(R2_, R3_));
    static constexpr int RL = rad(NST - 1);
    static constexpr int NRL = N_ / RL;
    static constexpr int PITCH = N_ + N_ / 8 + 1;  // contiguous layout: padded line pitch
    // strided tiles of G lines: with a power-of-two first radix, rows of G points plus one
    // spare row per 8 (its stride-R0 first-stage writes and stride-1 reads then fall in
    // distinct banks); otherwise rows of G + 1 points (an odd pitch spreads odd strides)
    static constexpr bool PAD8 = (R0_ & (R0_ - 1)) == 0;
    static constexpr int strided_lp(int G) { return PAD8 ? G : G + 1; }
    static constexpr int strided_tile(int G) { return PAD8 ? (N_ + N_ / 8 + 1) * G : N_ * (G + 1); }
    static_assert(R0_ * (R1_ > 1 ? R1_ : 1) * (R2_ > 1 ? R2_ : 1) * (R3_ > 1 ? R3_ : 1) == N_, "radices must multiply to N");
};

// Per-cell inputs of the FP32 source with the phase split exp(-i qC.u_a) =
// exp(-i qC.U) exp(-i g_a.P): e0 = exp(-i qC.U) is formed once per cell with the
// FP64-reduced sincos; per atom only g_a.P (a few radians) remains.
struct CellInputsF {
    float P0, P1, P2;
    float u0, u1, u2;
    cx<float> e0;
    float rg;
    int ph;
};

__device__ __forceinline__ void cell_load_f(const PassArgs<float>& a, uint32_t off, bool valid, CellInputsF& c) {
    c.ph = -1;
    c.rg = 1.0f;
    c.P0 = c.P1 = c.P2 = 0.0f;
    c.u0 = c.u1 = c.u2 = 0.0f;
    c.e0 = {1.0f, 0.0f};
    if (!valid) return;
    c.ph = a.phase[off];
    if (a.region) c.rg = a.region[off];
    c.P0 = a.P[off];
    c.P1 = a.P[a.ncell + off];
    c.P2 = a.P[(size_t)2 * a.ncell + off];
    const double U0 = a.U[off], U1 = a.U[a.ncell + off], U2 = a.U[(size_t)2 * a.ncell + off];
    c.u0 = (float)U0;
    c.u1 = (float)U1;
    c.u2 = (float)U2;
    float sn, cs;
    sincos_fp64_reduced(a.qC0 * U0 + a.qC1 * U1 + a.qC2 * U2, sn, cs);
    c.e0 = {cs, -sn};
}

// The same inputs in two steps: the loads of a cell (raw_load_f), then the FP64-reduced
// phase and the conversions (cell_prepare_f), so that a kernel can issue the loads of
// several cells before any arithmetic.
struct RawCellF {
    float P0, P1, P2;
    double U0, U1, U2;
    float rg;
    int ph;
};

// Per-lane base pointers of a lane's first cell; cell x of the lane is element x of each
// (64-bit pointer arithmetic, so constant cell offsets fold into the load instructions).
struct CellPtrsF {
    const float *P0, *P1, *P2;
    const double *U0, *U1, *U2;
    const signed char* ph;
    const float* rg;
};

__device__ __forceinline__ CellPtrsF cell_ptrs_f(const PassArgs<float>& a, size_t c0) {
    CellPtrsF q;
    q.P0 = a.P + c0;
    q.P1 = q.P0 + a.ncell;
    q.P2 = q.P1 + a.ncell;
    q.U0 = a.U + c0;
    q.U1 = q.U0 + a.ncell;
    q.U2 = q.U1 + a.ncell;
    q.ph = a.phase + c0;
    q.rg = a.region ? a.region + c0 : nullptr;
    return q;
}

__device__ __forceinline__ void raw_load_f(const CellPtrsF& q, int x, bool valid, RawCellF& c) {
    c.ph = -1;
    c.rg = 1.0f;
    c.P0 = c.P1 = c.P2 = 0.0f;
    c.U0 = c.U1 = c.U2 = 0.0;
    if (!valid) return;
    c.ph = q.ph[x];
    if (q.rg) c.rg = q.rg[x];
    c.P0 = q.P0[x];
    c.P1 = q.P1[x];
    c.P2 = q.P2[x];
    c.U0 = q.U0[x];
    c.U1 = q.U1[x];
    c.U2 = q.U2[x];
}

// u in memory-dim order (u0 = U[ax0], ...), for cell_source_m.
__device__ __forceinline__ void cell_prepare_f(const PassArgs<float>& a, const RawCellF& r, CellInputsF& c) {
    c.ph = r.ph;
    c.rg = r.rg;
    c.P0 = r.P0;
    c.P1 = r.P1;
    c.P2 = r.P2;
    const float ux = (float)r.U0, uy = (float)r.U1, uz = (float)r.U2;
    c.u0 = a.ax0 == 0 ? ux : (a.ax0 == 1 ? uy : uz);
    c.u1 = a.ax1 == 0 ? ux : (a.ax1 == 1 ? uy : uz);
    c.u2 = a.ax2 == 0 ? ux : (a.ax2 == 1 ? uy : uz);
    c.e0 = {1.0f, 0.0f};
    if (r.ph < 0) return;
    float sn, cs;
    sincos_fp64_reduced(a.qC0 * r.U0 + a.qC1 * r.U1 + a.qC2 * r.U2, sn, cs);
    c.e0 = {cs, -sn};
}

// Source of one cell for one atom, with the atom table's su rows and u already in
// memory-dim order (tab[2 + 3 m + k] = su[ax_m][k], u_m = U[ax_m]): the same values as
// cell_source_f without the per-atom axis selection.  One-hot phases take the
// straight-line path; fractional phases loop over the phases.
__device__ __forceinline__ void cell_source_m(const PassArgs<float>& a, const float* atomtab, const CellInputsF& ci,
                                              uint32_t off, int atom, cx<float>& E, cx<float>& V0, cx<float>& V1,
                                              cx<float>& V2) {
    E = V0 = V1 = V2 = {0.0f, 0.0f};
    if (ci.ph < 0) return;
    auto add = [&](const float* tab, float w) {
        const float u0 = ci.u0 + (tab[2] * ci.P0 + tab[3] * ci.P1 + tab[4] * ci.P2);
        const float u1 = ci.u1 + (tab[5] * ci.P0 + tab[6] * ci.P1 + tab[7] * ci.P2);
        const float u2 = ci.u2 + (tab[8] * ci.P0 + tab[9] * ci.P1 + tab[10] * ci.P2);
        float sn, cs;
        sincos_small(tab[11] * ci.P0 + tab[12] * ci.P1 + tab[13] * ci.P2, sn, cs);
        const cx<float> fw = {tab[0] * w, tab[1] * w};
        const cx<float> e = cmul(cmul(fw, ci.e0), cx<float>{cs, -sn});
        E = E + e;
        V0 = V0 + cscale(e, u0);
        V1 = V1 + cscale(e, u1);
        V2 = V2 + cscale(e, u2);
    };
    if (!a.fraction) {
        add(atomtab + (atom * a.n_phase + ci.ph) * kAtomTab, ci.rg);
        return;
    }
    for (int p = 0; p < a.n_phase; ++p) {
        const float w = ci.rg * a.fraction[(size_t)p * a.ncell + off];
        if (w == 0.0f) continue;
        add(atomtab + (atom * a.n_phase + p) * kAtomTab, w);
    }
}

// One-hot phases: the source of one cell for one atom as e = w f e0 e^{-i g.P} and the
// displacement u_m = u[ax_m] + su[ax_m] . P in memory-dim order (the fields are e and
// e u_m; the same values as cell_source_m).
__device__ __forceinline__ void cell_source_1h(const PassArgs<float>& a, const float* atomtab, const CellInputsF& ci,
                                               int atom, cx<float>& e, float& u0, float& u1, float& u2) {
    e = {0.0f, 0.0f};
    u0 = u1 = u2 = 0.0f;
    if (ci.ph < 0) return;
    const float* tab = atomtab + (atom * a.n_phase + ci.ph) * kAtomTab;
    u0 = ci.u0 + (tab[2] * ci.P0 + tab[3] * ci.P1 + tab[4] * ci.P2);
    u1 = ci.u1 + (tab[5] * ci.P0 + tab[6] * ci.P1 + tab[7] * ci.P2);
    u2 = ci.u2 + (tab[8] * ci.P0 + tab[9] * ci.P1 + tab[10] * ci.P2);
    float sn, cs;
    sincos_small(tab[11] * ci.P0 + tab[12] * ci.P1 + tab[13] * ci.P2, sn, cs);
    const cx<float> fw = {tab[0] * ci.rg, tab[1] * ci.rg};
    e = cmul(cmul(fw, ci.e0), cx<float>{cs, -sn});
}

// Source of one cell for one atom from CellInputsF (same values as cell_source).
__device__ __forceinline__ void cell_source_f(const PassArgs<float>& a, const float* atomtab, const CellInputsF& ci,
                                              uint32_t off, int atom, cx<float>& E, cx<float>& V0, cx<float>& V1,
                                              cx<float>& V2) {
    E = V0 = V1 = V2 = {0.0f, 0.0f};
    if (ci.ph < 0) return;
    const int p_lo = a.fraction ? 0 : ci.ph;
    const int p_hi = a.fraction ? a.n_phase : ci.ph + 1;
    for (int p = p_lo; p < p_hi; ++p) {
        float w = ci.rg;
        if (a.fraction) w *= a.fraction[(size_t)p * a.ncell + off];
        if (w == 0.0f) continue;
        const float* tab = atomtab + (atom * a.n_phase + p) * kAtomTab;
        const float ux = ci.u0 + (tab[2] * ci.P0 + tab[3] * ci.P1 + tab[4] * ci.P2);
        const float uy = ci.u1 + (tab[5] * ci.P0 + tab[6] * ci.P1 + tab[7] * ci.P2);
        const float uz = ci.u2 + (tab[8] * ci.P0 + tab[9] * ci.P1 + tab[10] * ci.P2);
        float sn, cs;
        sincos_small(tab[11] * ci.P0 + tab[12] * ci.P1 + tab[13] * ci.P2, sn, cs);
        // w f e0 e^{-i s}
        const cx<float> fw = {tab[0] * w, tab[1] * w};
        const cx<float> e = cmul(cmul(fw, ci.e0), cx<float>{cs, -sn});
        E = E + e;
        V0 = V0 + cscale(e, a.ax0 == 0 ? ux : (a.ax0 == 1 ? uy : uz));
        V1 = V1 + cscale(e, a.ax1 == 0 ? ux : (a.ax1 == 1 ? uy : uz));
        V2 = V2 + cscale(e, a.ax2 == 0 ? ux : (a.ax2 == 1 ? uy : uz));
    }
}

// The source inputs of an FP64 kernel on FP32 arrays seen as an FP32 pass's arguments (the
// FP32 source helpers read P, U, the phases, the weights, qC, the axes and n_phase only).
template <typename T>
__device__ __forceinline__ PassArgs<float> source_view(const PassArgs<T>& a) {
    if constexpr (std::is_same<T, float>::value) {
        return a;
    } else {
        PassArgs<float> f{};
        f.P = reinterpret_cast<const float*>(a.P);
        f.U = a.U;
        f.phase = a.phase;
        f.fraction = reinterpret_cast<const float*>(a.fraction);
        f.region = reinterpret_cast<const float*>(a.region);
        f.ncell = a.ncell;
        f.qC0 = a.qC0;
        f.qC1 = a.qC1;
        f.qC2 = a.qC2;
        f.ax0 = a.ax0;
        f.ax1 = a.ax1;
        f.ax2 = a.ax2;
        f.n_phase = a.n_phase;
        return f;
    }
}

// Workgroup barrier that orders LDS only.  __syncthreads() is a release/acquire fence
// on every address space, so the compiler waits for all outstanding global loads
// (vmcnt 0) before each barrier, which ends a register prefetch at the first LDS
// hand-off.  These kernels exchange only LDS data across barriers: every global load
// whose value a barrier publishes feeds an LDS write first, and the loads left in
// flight (the next input line) never alias the stores made meanwhile (distinct
// fields or buffers), so the fence is narrowed to the LDS address space.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The same ordering of LDS accesses without the memory-model fence, whose release part also
// waits for LDS writes of global_load_lds still in flight (vmcnt): this wave's LDS accesses
// complete (lgkmcnt(0)), then the workgroup barrier; the memory clobber keeps the compiler
// from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier_nofence() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// s_waitcnt immediate that waits until at most n vector-memory operations are outstanding
// (expcnt and lgkmcnt left at their maxima, which wait for nothing)
constexpr int vmcnt_imm(int n) {
    return (n & 0xF) | (0x7 << 4) | (0xF << 8) | ((n >> 4) << 14);
}

template <bool NOFENCE>
__device__ __forceinline__ void stage_barrier() {
    if constexpr (NOFENCE)
        lds_barrier_nofence();
    else
        lds_barrier();
}

// p[byte_off] with p wave-uniform and byte_off a 32-bit lane offset, as a buffer access: a
// resource (scalar registers) over the 4 GiB from p and the lane offset as its 32-bit
// vector operand.  A plain p + byte_off leaves the optimiser free to regroup the sum as
// (row offset + lane offset) + base and keep a 64-bit vector address per row, and to
// widen the offset outside a loop, after which the loop's accesses cannot use the 32-bit
// form; the buffer form fixes the operands.  The resource's last word is untyped data
// (format 32); the range check (offset < 2^32 - 1) never triggers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0xFFFFFFFF, 0x00020000);
}
template <typename V>
__device__ __forceinline__ V load_at(const V* p, uint32_t byte_off) {
    static_assert(sizeof(V) == 4 || sizeof(V) == 8 || sizeof(V) == 16, "4, 8 or 16 bytes");
    if constexpr (sizeof(V) == 4)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(rsrc_at(p), byte_off, 0, 0));
    else if constexpr (sizeof(V) == 8)
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(rsrc_at(p), byte_off, 0, 0));
    else
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(rsrc_at(p), byte_off, 0, 0));
}
template <typename V>
__device__ __forceinline__ void store_at(V* p, uint32_t byte_off, const V& v) {
    static_assert(sizeof(V) == 8 || sizeof(V) == 16, "8 or 16 bytes");
    if constexpr (sizeof(V) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v),
                                              rsrc_at(p), byte_off, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v),
                                               rsrc_at(p), byte_off, 0, 0);
}

template <bool STRIDED, class SZ, int LP>
__device__ __forceinline__ int lds_addr(int line, int idx) {
    if constexpr (STRIDED)
        return (SZ::PAD8 ? idx + (idx >> 3) : idx) * LP + line;
    else
        return line * SZ::PITCH + idx + (idx >> 3);
}

// Layout of the per-stage twiddle table: stage s starts at off(s) with ntw(s) twiddles
// W_N^{r k N / (NS R)} at (r - 1) NS + k.
template <class SZ>
struct StageTw {
    static constexpr int ntw(int s) { return SZ::ns(s) > 1 ? (SZ::rad(s) - 1) * SZ::ns(s) : 0; }
    static constexpr int off(int s) { return s == 0 ? 0 : off(s - 1) + ntw(s - 1); }
    static constexpr int total() { return off(SZ::NST); }
};

template <typename T, class SZ, int S>
__device__ __forceinline__ void fill_stage_tw(cx<T>* tw, const cx<T>* __restrict__ W, int tid, int nthr) {
    if constexpr (S < SZ::NST) {
        constexpr int R = SZ::rad(S), NS = SZ::ns(S);
        constexpr int OFF = StageTw<SZ>::off(S), NTW = StageTw<SZ>::ntw(S);
        constexpr int TWS = SZ::N / (NS * R);
        for (int e = tid; e < NTW; e += nthr) tw[OFF + e] = W[(e / NS + 1) * (e % NS) * TWS];
    }
}

// Stage s on F lines held in registers v[f][r] (inputs of butterfly j already loaded):
// twiddle (shared by the F lines), then DFT_R.
template <typename T, class SZ, int S, int F>
__device__ __forceinline__ void stage_compute(cx<T> (&v)[F][SZ::RMAX], int j, const cx<T>* __restrict__ W) {
    constexpr int R = SZ::rad(S);
    constexpr int NS = SZ::ns(S);
    if constexpr (NS > 1) {
        constexpr int OFF = StageTw<SZ>::off(S);
        const int k = j % NS;
        cx<T> w[R];
#pragma unroll
        for (int r = 1; r < R; ++r) w[r] = W[OFF + (r - 1) * NS + k];
#pragma unroll
        for (int f = 0; f < F; ++f)
#pragma unroll
            for (int r = 1; r < R; ++r) v[f][r] = cmul(v[f][r], w[r]);
    }
#pragma unroll
    for (int f = 0; f < F; ++f) {
        cx<T> t[R];
#pragma unroll
        for (int r = 0; r < R; ++r) t[r] = v[f][r];
        dft<R>(t);
#pragma unroll
        for (int r = 0; r < R; ++r) v[f][r] = t[r];
    }
}

// Stage S (odd radix R >= 5, not the last) of one line held in v: twiddle, then the DFT_R
// outputs written to their Stockham positions in LDS a symmetric pair at a time, so that the R
// outputs are never all held in registers beside the pair sums (stage_compute + stage_write
// hold both: about 2 R complex values).
template <typename T, class SZ, int S, bool STRIDED, int LP>
__device__ __forceinline__ void stage_odd_to_lds(cx<T>* tile, cx<T> (&v)[1][SZ::RMAX], int j, int line,
                                                 const cx<T>* __restrict__ W) {
    constexpr int R = SZ::rad(S), NS = SZ::ns(S), H = (R - 1) / 2;
    if constexpr (NS > 1) {
        constexpr int OFF = StageTw<SZ>::off(S);
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) v[0][r] = cmul(v[0][r], W[OFF + (r - 1) * NS + k]);
    }
    cx<T> a[H + 1], b[H + 1];
    cx<T> y0 = v[0][0];
#pragma unroll
    for (int s = 1; s <= H; ++s) {
        a[s] = v[0][s] + v[0][R - s];
        b[s] = v[0][s] - v[0][R - s];
        y0 = y0 + a[s];
    }
    const int o = (j / NS) * NS * R + (j % NS);
    tile[lds_addr<STRIDED, SZ, LP>(line, o)] = y0;
    constexpr bool f32 = std::is_same<T, float>::value;
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        cx<T> A = v[0][0];
        cx<T> B = {(T)0, (T)0};
#pragma unroll
        for (int s = 1; s <= H; ++s) {
            const int m = (t * s) % R;
            const int mm = m <= H ? m : R - m;
            const T c = f32 ? (T)kOddRoots<R>.c[mm] : (T)kOddRoots64<R>.c[mm];
            const T s1 = f32 ? (T)kOddRoots<R>.s[mm] : (T)kOddRoots64<R>.s[mm];
            const T sn = m <= H ? s1 : -s1;
            A.x += a[s].x * c;
            A.y += a[s].y * c;
            B.x += b[s].x * sn;
            B.y += b[s].y * sn;
        }
        tile[lds_addr<STRIDED, SZ, LP>(line, o + t * NS)] = {A.x + B.y, A.y - B.x};
        tile[lds_addr<STRIDED, SZ, LP>(line, o + (R - t) * NS)] = {A.x - B.y, A.y + B.x};
    }
}

// Write the outputs of stage S (butterfly j) to LDS at their Stockham positions.
template <typename T, class SZ, int S, int F, bool STRIDED, int LP>
__device__ __forceinline__ void stage_write(cx<T>* tile, const cx<T> (&v)[F][SZ::RMAX], int j, int line0, int dline) {
    constexpr int R = SZ::rad(S), NS = SZ::ns(S);
    const int o = (j / NS) * NS * R + (j % NS);
#pragma unroll
    for (int f = 0; f < F; ++f)
#pragma unroll
        for (int r = 0; r < R; ++r)
            tile[lds_addr<STRIDED, SZ, LP>(line0 + f * dline, o + r * NS)] = v[f][r];
}

// Read the inputs x[j + r N/R_S] of stage S from LDS.
template <typename T, class SZ, int S, int F, bool STRIDED, int LP>
__device__ __forceinline__ void stage_read(const cx<T>* tile, cx<T> (&v)[F][SZ::RMAX], int j, int line0, int dline) {
    constexpr int R = SZ::rad(S), NR = SZ::nr(S);
#pragma unroll
    for (int f = 0; f < F; ++f)
#pragma unroll
        for (int r = 0; r < R; ++r) v[f][r] = tile[lds_addr<STRIDED, SZ, LP>(line0 + f * dline, j + r * NR)];
}

// Stages 0 .. NST-1 of the F lines of butterfly j, with two LDS buffers X, Y for the
// hand-offs between stages.  LOAD0: the stage-0 inputs are in X in natural order
// (written before a barrier); otherwise they are already in v.  Stage outputs go
// to the buffers in turn (Y, X, Y, ... with LOAD0; X, Y, X, ... without), with a
// barrier between each write and the read that follows; the last stage leaves
// X[j + r NRL] in v.  No barrier is pending on return.  The next call may write
// the buffer this call read last only after a barrier, so callers alternate the
// roles of X and Y between calls when the number of writes (chunk_writes) is odd.
// ONEBUF (LOAD0 false): X is Y, and a barrier also precedes every write, after the
// reads of the same buffer before it (this call's or the previous call's); half the
// LDS at twice the barriers.
template <class SZ, bool LOAD0>
constexpr int chunk_writes() {
    return SZ::NST - 1 + (LOAD0 ? 1 : 0);
}

// FIRSTBAR false (ONEBUF): the caller has already put a barrier between the previous reads of
// the buffer and this call, so the one before the first stage's write is left out.
template <typename T, class SZ, int F, bool STRIDED, int LP, bool LOAD0, bool ONEBUF = false, bool NOFENCE = false,
          bool FIRSTBAR = true>
__device__ __forceinline__ void transform(cx<T>* X, cx<T>* Y, cx<T> (&v)[F][SZ::RMAX], int j, int line0, int dline,
                                          const cx<T>* __restrict__ W) {
    static_assert(!(ONEBUF && LOAD0), "one buffer holds the stage-0 inputs only when they are in registers");
    cx<T>* rd = X;  // buffer holding the current stage's inputs
    cx<T>* wr = LOAD0 ? Y : X;
    const bool a0 = j < SZ::nr(0);
    if constexpr (LOAD0) {
        if (a0) stage_read<T, SZ, 0, F, STRIDED, LP>(rd, v, j, line0, dline);
    }
    // an odd first radix (>= 5) of a multi-stage line writes its outputs as it forms them
    constexpr bool ODD0 = F == 1 && SZ::NST > 1 && (SZ::rad(0) & 1) == 1 && SZ::rad(0) >= 5;
    if constexpr (ODD0) {
        if constexpr (ONEBUF && FIRSTBAR) stage_barrier<NOFENCE>();
        if (a0) stage_odd_to_lds<T, SZ, 0, STRIDED, LP>(wr, v, j, line0, W);
    } else {
        if (a0) stage_compute<T, SZ, 0, F>(v, j, W);
    }
    if constexpr (SZ::NST > 1) {
        if constexpr (!ODD0) {
            if constexpr (ONEBUF && FIRSTBAR) stage_barrier<NOFENCE>();
            if (a0) stage_write<T, SZ, 0, F, STRIDED, LP>(wr, v, j, line0, dline);
        }
        stage_barrier<NOFENCE>();
        rd = wr;
        wr = (wr == X) ? Y : X;
        const bool a1 = j < SZ::nr(1);
        if (a1) {
            stage_read<T, SZ, 1, F, STRIDED, LP>(rd, v, j, line0, dline);
            stage_compute<T, SZ, 1, F>(v, j, W);
        }
        if constexpr (SZ::NST > 2) {
            if constexpr (ONEBUF) {
                stage_barrier<NOFENCE>();
                wr = X;
            }
            if (a1) stage_write<T, SZ, 1, F, STRIDED, LP>(wr, v, j, line0, dline);
            stage_barrier<NOFENCE>();
            rd = wr;
            wr = (wr == X) ? Y : X;
            const bool a2 = j < SZ::nr(2);
            if (a2) {
                stage_read<T, SZ, 2, F, STRIDED, LP>(rd, v, j, line0, dline);
                stage_compute<T, SZ, 2, F>(v, j, W);
            }
            if constexpr (SZ::NST > 3) {
                if constexpr (ONEBUF) {
                    stage_barrier<NOFENCE>();
                    wr = X;
                }
                if (a2) stage_write<T, SZ, 2, F, STRIDED, LP>(wr, v, j, line0, dline);
                stage_barrier<NOFENCE>();
                rd = wr;
                const bool a3 = j < SZ::nr(3);
                if (a3) {
                    stage_read<T, SZ, 3, F, STRIDED, LP>(rd, v, j, line0, dline);
                    stage_compute<T, SZ, 3, F>(v, j, W);
                }
            }
        }
    }
}

// A last stage of prime radix R shared by P lanes per butterfly (the last stage of a line
// with few butterflies leaves most lanes idle: 68 = 4 x 17 has four radix-17 butterflies on
// seventeen lanes).  With g a primitive root mod R, the P cosets of the subgroup <g^P> split
// the outputs 1..R-1; lane q of a butterfly computes outputs g^q t0 and -g^q t0 for the
// representatives t0 of that subgroup (up to sign), so that every lane runs the same
// constants: X[g^q t0] = sum_s x[g^-q s] w^(t0 s), i.e. the lane reads its inputs in the
// order r = g^-q s mod R and forms the standard symmetric-pair sums of X[t0].  Lane 0 also
// forms X[0].  (R - 1) / P must be even, so that the subgroup holds -1.
template <int R, int P>
struct SplitOdd {
    static constexpr int H = (R - 1) / 2;
    static constexpr int PPL = (R - 1) / (2 * P);  // output pairs a lane
    static constexpr int SLOTS = 2 * PPL + 1;      // X[t0], X[-t0] per pair, then X[0] (lane 0)
    static constexpr int pw(int b, int e) { return e == 0 ? 1 : (pw(b, e - 1) * b) % R; }
    static constexpr bool primitive(int g) {
        for (int e = 1; e < R - 1; ++e)
            if (pw(g, e) == 1) return false;
        return true;
    }
    static constexpr int root() {
        for (int g = 2; g < R; ++g)
            if (primitive(g)) return g;
        return 0;
    }
    static constexpr int G = root();
    static constexpr int GINV = pw(G, R - 2);
    // the subgroup's representatives up to sign (the members at most H)
    static constexpr int t0(int i) {
        int n = 0;
        for (int e = 0; e < (R - 1) / P; ++e) {
            const int m = pw(G, P * e);
            if (m <= H) {
                if (n == i) return m;
                ++n;
            }
        }
        return 0;
    }
    static_assert(R >= 5 && (R - 1) % (2 * P) == 0, "the subgroup must hold -1");
};

// The representatives t0(k) and the powers g^q of a split as a compile-time table (indexed by
// unrolled loop counters, its entries fold into the code).
template <int R, int P>
struct SplitTab {
    int t0[SplitOdd<R, P>::PPL];
    int gq[P];
    constexpr SplitTab() : t0(), gq() {
        for (int k = 0; k < SplitOdd<R, P>::PPL; ++k) t0[k] = SplitOdd<R, P>::t0(k);
        for (int q = 0; q < P; ++q) gq[q] = SplitOdd<R, P>::pw(SplitOdd<R, P>::G, q);
    }
};
template <int R, int P>
__device__ constexpr SplitTab<R, P> kSplitTab{};

// Slot i of lane q (q < P): its output index t in 0..R-1 (slot 2k: g^q t0(k), 2k+1: -g^q t0(k),
// 2 PPL: 0); -1 where the lane has no such output (X[0] on lanes q > 0).
template <int R, int P>
__device__ __forceinline__ int split_slot_index(int q, int i) {
    using SO = SplitOdd<R, P>;
    if (i == 2 * SO::PPL) return q == 0 ? 0 : -1;
    int gq = 1;
#pragma unroll
    for (int e = 1; e < P; ++e) gq = q == e ? kSplitTab<R, P>.gq[e] : gq;
    const int t = (gq * kSplitTab<R, P>.t0[i >> 1]) % R;
    return (i & 1) ? R - t : t;
}

// Slots 2K, 2K + 1 (X[t0], X[-t0] with t0 = t0(K), a compile-time constant) and the pairs after
// them, from the symmetric sums av, bv of the lane's inputs v.
template <typename T, int R, int P, int K>
__device__ __forceinline__ void split_pairs(cx<T>* v, const cx<T>* av, const cx<T>* bv, std::integral_constant<int, K>) {
    using SO = SplitOdd<R, P>;
    if constexpr (K < SO::PPL) {
        constexpr int H = SO::H, t = SO::t0(K);
        constexpr bool f32 = std::is_same<T, float>::value;
        cx<T> A = v[0];
        cx<T> B = {(T)0, (T)0};
#pragma unroll
        for (int s = 1; s <= H; ++s) {
            const int m = (t * s) % R;
            const int mm = m <= H ? m : R - m;
            const T c = f32 ? (T)kOddRoots<R>.c[mm] : (T)kOddRoots64<R>.c[mm];
            const T s1 = f32 ? (T)kOddRoots<R>.s[mm] : (T)kOddRoots64<R>.s[mm];
            const T sn = m <= H ? s1 : -s1;
            A.x += av[s].x * c;
            A.y += av[s].y * c;
            B.x += bv[s].x * sn;
            B.y += bv[s].y * sn;
        }
        // v[0] is read by every pair: the later pairs are formed before this one is stored
        split_pairs<T, R, P>(v, av, bv, std::integral_constant<int, K + 1>{});
        v[2 * K] = {A.x + B.y, A.y - B.x};
        v[2 * K + 1] = {A.x - B.y, A.y + B.x};
    }
}

// The slots of one lane of a split radix-R DFT whose inputs v[0..R) are already in the lane's
// order (x[g^-q s] in v[s]): X[g^q t0(k)], X[-g^q t0(k)] in v[2k], v[2k + 1], then X[0] in
// v[2 PPL] (used by lane 0 only).
template <typename T, int R, int P, int N>
__device__ __forceinline__ void dft_split_slots(cx<T> (&v)[N]) {
    using SO = SplitOdd<R, P>;
    constexpr int H = SO::H;
    cx<T> av[H + 1], bv[H + 1];
    cx<T> y0 = v[0];
#pragma unroll
    for (int s = 1; s <= H; ++s) {
        av[s] = v[s] + v[R - s];
        bv[s] = v[s] - v[R - s];
        y0 = y0 + av[s];
    }
    split_pairs<T, R, P>(v, av, bv, std::integral_constant<int, 0>{});
    v[2 * SO::PPL] = y0;
}

// The last stage of a line, split over P lanes per butterfly: lane (b, q) = (j % NRL, j / NRL)
// reads the R inputs of butterfly b in its order, twiddles them and leaves its slots in v.
template <typename T, class SZ, int P>
__device__ __forceinline__ void last_stage_split(const cx<T>* line, cx<T> (&v)[SZ::RMAX], int b, int q,
                                                 const cx<T>* tw) {
    constexpr int L = SZ::NST - 1, R = SZ::rad(L), NR = SZ::nr(L), NS = SZ::ns(L);
    constexpr int OFF = StageTw<SZ>::off(L);
    using SO = SplitOdd<R, P>;
    int gi = 1;  // g^-q mod R
#pragma unroll
    for (int e = 0; e + 1 < P; ++e) gi = e < q ? (gi * SO::GINV) % R : gi;
    v[0] = line[lds_addr<false, SZ, 0>(0, b)];
    int r = 0;
#pragma unroll
    for (int s = 1; s < R; ++s) {
        r += gi;
        r = r >= R ? r - R : r;  // r = g^-q s mod R
        const cx<T> x = line[lds_addr<false, SZ, 0>(0, b + r * NR)];
        v[s] = NS > 1 ? cmul(x, tw[OFF + (r - 1) * NS + b]) : x;
    }
    dft_split_slots<T, R, P>(v);
}

// Outputs a lane holds after a line transform, at the row's point index
// out_index(j, i) for slot i < out_slots (i.e. X[j + i N/RL] when the last stage is not split).
template <class SZ, int P>
struct LineOut {
    static constexpr int L = SZ::NST - 1;
    static constexpr int NRL = SZ::NRL;
    static constexpr int slots() {
        if constexpr (P > 1)
            return SplitOdd<SZ::rad(L), P>::SLOTS;
        else
            return SZ::RL;
    }
    static constexpr int SLOTS = slots();
    // the lane holds outputs at all
    __device__ static bool owns(int j) { return j < P * NRL; }
    // point index of slot i, or -1
    __device__ static int index(int j, int i) {
        if constexpr (P > 1) {
            const int t = split_slot_index<SZ::rad(L), P>(j / NRL, i);
            return t < 0 ? -1 : j % NRL + t * NRL;
        } else {
            return j + i * NRL;
        }
    }
};

// transform() for nlines contiguous lines whose stage-0 inputs are in X in natural order
// (written before a barrier), one line per thread group: the stages before the last
// run with their butterflies packed over the block's threads (thread t takes butterfly
// t mod nr(s) of line t / nr(s)), so a stage with fewer butterflies than threads per line
// (a large radix after a small one's worth of threads) keeps whole waves busy and leaves
// the rest idle instead of running every wave with a few active lanes; the last stage
// runs in the thread's own line (line, j) and leaves X[j + r NRL] in v.  Same buffers,
// writes and barriers as transform<..., LOAD0 = true>.
// Stage 0 of transform_packed split over P0 threads per butterfly (P0 > 1; empty otherwise).
template <typename T, class SZ, int P0>
__device__ __forceinline__ void packed_stage0_split(const cx<T>* rd, cx<T>* wr, cx<T> (&v)[1][SZ::RMAX], int tid,
                                                int nlines) {
    if constexpr (P0 > 1) {
        constexpr int NR = SZ::nr(0);
        constexpr int R0 = SZ::rad(0);
        using SO = SplitOdd<R0, P0>;
        int ln = tid / (NR * P0);
        const int rem = tid - ln * (NR * P0);
        int b = rem % NR, q = rem / NR;
        // the line, butterfly and coset are re-formed for every use (the empty asm hides their
        // values), so that the compiler does not hold this stage's eighteen-odd LDS addresses
        // across the chunks
        asm volatile("" : "+v"(ln), "+v"(b), "+v"(q));
        if (ln < nlines) {
            int gi = 1;  // g^-q mod R0
#pragma unroll
            for (int e = 0; e + 1 < P0; ++e) gi = e < q ? (gi * SO::GINV) % R0 : gi;
            v[0][0] = rd[lds_addr<false, SZ, 0>(ln, b)];
            int r = 0;
#pragma unroll
            for (int s = 1; s < R0; ++s) {
                r += gi;
                r = r >= R0 ? r - R0 : r;
                v[0][s] = rd[lds_addr<false, SZ, 0>(ln, b + r * NR)];
            }
            dft_split_slots<T, R0, P0>(v[0]);
#pragma unroll
            for (int i = 0; i < SO::SLOTS; ++i) {
                const int t = split_slot_index<R0, P0>(q, i);
                if (t >= 0) wr[lds_addr<false, SZ, 0>(ln, b * R0 + t)] = v[0][i];
            }
        }
    }
}

// P0 > 1: stage 0 (prime radix R0, packed) runs P0 threads per butterfly, thread t taking coset
// q = (t mod NR0 P0) / NR0 of butterfly b = t mod NR0 of line t / (NR0 P0) (see SplitOdd): it
// reads the butterfly's inputs in the coset's order and writes its slots to their outputs.
template <typename T, class SZ, int P0 = 1>
__device__ __forceinline__ void transform_packed(cx<T>* X, cx<T>* Y, cx<T> (&v)[1][SZ::RMAX], int tid, int nlines,
                                                 int j, int line, const cx<T>* __restrict__ W) {
    constexpr int SL = SZ::NST - 1;
    cx<T>* rd = X;
    cx<T>* wr = Y;
    auto packed_stage = [&](auto s_) {
        constexpr int S = decltype(s_)::value;
        constexpr int NR = SZ::nr(S);
        if constexpr (S == 0 && P0 > 1) {
            packed_stage0_split<T, SZ, P0>(rd, wr, v, tid, nlines);
        } else {
            const int ln = tid / NR, jj = tid - ln * NR;
            if (ln < nlines) {
                stage_read<T, SZ, S, 1, false, 0>(rd, v, jj, ln, 0);
                stage_compute<T, SZ, S, 1>(v, jj, W);
                stage_write<T, SZ, S, 1, false, 0>(wr, v, jj, ln, 0);
            }
        }
        lds_barrier();
        cx<T>* t = rd;
        rd = wr;
        wr = t;
    };
    if constexpr (SL > 0) packed_stage(std::integral_constant<int, 0>{});
    if constexpr (SL > 1) packed_stage(std::integral_constant<int, 1>{});
    if constexpr (SL > 2) packed_stage(std::integral_constant<int, 2>{});
    if (j < SZ::nr(SL)) {
        stage_read<T, SZ, SL, 1, false, 0>(rd, v, j, line, 0);
        stage_compute<T, SZ, SL, 1>(v, j, W);
    }
}

// Fill the per-stage twiddle table (StageTw layout, N - R0 entries at most) in LDS from the
// table W[m] = exp(-2 pi i m / N) (caller synchronises before use).  Entry (r - 1) NS + k of
// stage s holds the twiddle of input r of butterfly k, so the lanes of a wave, which take
// consecutive k, read consecutive entries (W[r k N / (NS R)] itself is strided by r N / (NS R)
// entries between them, up to 16-way LDS bank conflicts for complex FP64 values).
template <typename T, class SZ>
__device__ __forceinline__ void load_twiddles(cx<T>* dst, const cx<T>* __restrict__ src, int tid, int nthr) {
    fill_stage_tw<T, SZ, 1>(dst, src, tid, nthr);
    fill_stage_tw<T, SZ, 2>(dst, src, tid, nthr);
    fill_stage_tw<T, SZ, 3>(dst, src, tid, nthr);
}

// Multiply the last-stage outputs by the chunk coefficient and add them to the
// group's outputs (input 0 -> out 0, input 1 -> out 0 times -i k, input f -> out f-1).
template <typename T, class SZ, int F, int NOUT>
__device__ __forceinline__ void accumulate_regs(cx<T> (&acc)[NOUT][SZ::RL], const cx<T> (&v)[F][SZ::RMAX], int j,
                                                const PassArgs<T>& a, int ct) {
#pragma unroll
    for (int r = 0; r < SZ::RL; ++r) {
        const int idx = j + r * SZ::NRL;
        const cx<T> cf = chunk_coef(a, ct, idx);
        const cx<T> t0 = cmul(cf, v[0][r]);
        acc[0][r] = acc[0][r] + t0;
        if constexpr (F > 1) {
            const T kv = a.k[idx];
            acc[0][r] = acc[0][r] + cscale(negi(cmul(cf, v[1][r])), kv);
        }
#pragma unroll
        for (int f = 2; f < F; ++f) acc[f - 1][r] = acc[f - 1][r] + cmul(cf, v[f][r]);
    }
}

// ---------------------------------------------------------------------------
// Contiguous pass: block = G rows of N points; thread (row g = tid / TPL, j = tid % TPL).
// SOURCE: the F = 4 inputs of a chunk are the source values of one atom; otherwise
// they are F consecutive input fields (F = nin).
template <typename T, class SZ, int G, int F, bool SOURCE>
__global__ __launch_bounds__(G* SZ::TPL) void fast_contig(PassArgs<T> a) {
    static_assert(!SOURCE, "the source pass uses fast_source");
    constexpr int N = SZ::N, TPL = SZ::TPL, R0 = SZ::rad(0), NR0 = SZ::nr(0);
    constexpr int NOUT = F > 1 ? F - 1 : 1;
    constexpr int TILE = F * G * SZ::PITCH;
    constexpr bool ALT = (chunk_writes<SZ, false>() & 1) == 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* bufA = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* bufB = bufA + TILE;
    cx<T>* Wl = bufB + TILE;
    const int tid = threadIdx.x;
    const int g = tid / TPL, j = tid % TPL;
    const uint32_t nrows = a.n1 * a.n2;
    const uint32_t row = blockIdx.x * G + g;
    const bool row_ok = row < nrows;
    const uint32_t cell0 = row * (uint32_t)N;
    load_twiddles<T, SZ>(Wl, a.W, tid, G * TPL);
    lds_barrier();
    int par = 0;
    for (int grp = 0; grp < a.ngroups; ++grp) {
        cx<T> acc[NOUT][SZ::RL];
#pragma unroll
        for (int o = 0; o < NOUT; ++o)
#pragma unroll
            for (int r = 0; r < SZ::RL; ++r) acc[o][r] = {(T)0, (T)0};
        for (int c = uload(a.group_first, grp); c < uload(a.group_first, grp + 1); ++c) {
            const int src = uload(a.chunk_src, c);
            const int ct = uload(a.chunk_coef, c);
            cx<T> v[F][SZ::RMAX];
            if (j < NR0) {
                const cx<T>* in = a.in + (size_t)src * a.nin * a.ncell + cell0 + j;
#pragma unroll
                for (int r = 0; r < R0; ++r)
#pragma unroll
                    for (int f = 0; f < F; ++f)
                        v[f][r] = row_ok ? in[(size_t)f * a.ncell + r * NR0] : cx<T>{(T)0, (T)0};
            }
            transform<T, SZ, F, false, 0, false>(par ? bufB : bufA, par ? bufA : bufB, v, j, g, G, Wl);
            if (ALT) par ^= 1;
            if (j < SZ::NRL) accumulate_regs<T, SZ, F, NOUT>(acc, v, j, a, ct);
        }
        if (row_ok && j < SZ::NRL) {
            cx<T>* o = a.out + (size_t)grp * NOUT * a.ncell + cell0 + j;
#pragma unroll
            for (int q = 0; q < NOUT; ++q)
#pragma unroll
                for (int r = 0; r < SZ::RL; ++r) o[(size_t)q * a.ncell + r * SZ::NRL] = acc[q][r];
        }
    }
}

// ---------------------------------------------------------------------------
// Contiguous pass with the source (pass 0): block = G rows of N cells, 4 * G * TPL
// threads.  For each atom the block first computes the four source values of its
// cells (about R0 / 4 cells per thread, inputs cached in registers) into the tile
// in natural order; then thread (f, row, j) transforms line (f, row) and keeps the
// contribution of input f to its output: f = 0 and f = 1 (times -i k) both feed
// output 0 and are summed through LDS once per group; f = 2, 3 feed outputs 1, 2.
// IO32 (FP64 only): FP64 arithmetic on FP32 P, weights and outputs.
// P0 > 1: stage 0 split over P0 threads per butterfly (transform_packed).
template <typename T, class SZ, int G, bool IO32 = false, int P0 = 1>
__global__ __launch_bounds__(4 * G * SZ::TPL) void fast_source(PassArgs<T> a) {
    using TS = io_t<T, IO32>;
    constexpr int N = SZ::N, TPL = SZ::TPL, NRL = SZ::NRL, RL = SZ::RL;
    constexpr int NTHR = 4 * G * TPL;
    constexpr int CPT = (G * N + NTHR - 1) / NTHR;  // cells per thread in the source phase
    constexpr int TILE = 4 * G * SZ::PITCH;
    constexpr bool ALT = (chunk_writes<SZ, true>() & 1) == 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* bufA = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* bufB = bufA + TILE;
    cx<T>* Wl = bufB + TILE;
    // FP32 source (FP32 kernels, and FP64 ones on FP32 arrays): the FP32 atom table
    constexpr bool F32 = std::is_same<T, float>::value;
    constexpr bool SRC32 = F32 || IO32;
    using TT = std::conditional_t<SRC32, float, T>;
    TT* tabl = reinterpret_cast<TT*>(Wl + N);  // atom table (n_atom * n_phase * kAtomTab)
    const int tid = threadIdx.x;
    const int f = tid / (G * TPL);
    const int g = (tid / TPL) % G;
    const int j = tid % TPL;
    const uint32_t nrows = a.n1 * a.n2;
    const uint32_t row0 = blockIdx.x * G;
    const uint32_t cell0 = row0 * (uint32_t)N;
    const uint32_t row = row0 + g;
    load_twiddles<T, SZ>(Wl, a.W, tid, NTHR);
    const int ntab = a.n_atom_tab;
    for (int m = tid; m < ntab; m += NTHR) tabl[m] = (TT)a.atomtab[m];

    // source cells of this thread: block cell e = tid + m * NTHR -> (row e / N, x e % N);
    // the FP32 source splits the phase (CellInputsF) and is widened as it enters the FP64
    // transforms of an FP64 kernel on FP32 arrays (FP64 phase reduction, FP32 values: the
    // rounding is relative per cell, like the storage's); FP64 plans form it whole per atom
    // (cell_source)
    const PassArgs<float> af = source_view(a);
    using CellIn = std::conditional_t<SRC32, CellInputsF, CellInputs<T>>;
    CellIn ci[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int e = tid + m * NTHR;
        const bool valid = e < G * N && row0 + e / N < nrows;
        if constexpr (SRC32)
            cell_load_f(af, cell0 + e, valid, ci[m]);
        else
            cell_load<T, IO32>(a, cell0 + e, valid, ci[m]);
    }
    lds_barrier();

    int par = 0;
    for (int grp = 0; grp < a.ngroups; ++grp) {
        cx<T> acc[RL];
#pragma unroll
        for (int r = 0; r < RL; ++r) acc[r] = {(T)0, (T)0};
        for (int c = uload(a.group_first, grp); c < uload(a.group_first, grp + 1); ++c) {
            const int src = uload(a.chunk_src, c);
            const int ct = uload(a.chunk_coef, c);
            cx<T>* X = par ? bufB : bufA;
            cx<T>* Y = par ? bufA : bufB;
            if (ALT) par ^= 1;
#pragma unroll
            for (int m = 0; m < CPT; ++m) {
                const int e = tid + m * NTHR;
                if (e < G * N) {
                    cx<T> v0, v1, v2, v3;
                    if constexpr (SRC32) {
                        cx<float> s0, s1, s2, s3;
                        cell_source_f(af, tabl, ci[m], cell0 + e, src, s0, s1, s2, s3);
                        v0 = cx_as<T>(s0);
                        v1 = cx_as<T>(s1);
                        v2 = cx_as<T>(s2);
                        v3 = cx_as<T>(s3);
                    } else {
                        cell_source<T, IO32>(a, tabl, ci[m], cell0 + e, src, v0, v1, v2, v3);
                    }
                    const int rr = e / N, x = e - rr * N;
                    X[lds_addr<false, SZ, 0>(0 * G + rr, x)] = v0;
                    X[lds_addr<false, SZ, 0>(1 * G + rr, x)] = v1;
                    X[lds_addr<false, SZ, 0>(2 * G + rr, x)] = v2;
                    X[lds_addr<false, SZ, 0>(3 * G + rr, x)] = v3;
                }
            }
            lds_barrier();
            cx<T> v[1][SZ::RMAX];
            transform_packed<T, SZ, P0>(X, Y, v, tid, 4 * G, j, f * G + g, Wl);
            if (j < NRL) {
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int idx = j + r * NRL;
                    cx<T> t = cmul(chunk_coef(a, ct, idx), v[0][r]);
                    if (f == 1) t = cscale(negi(t), a.k[idx]);
                    acc[r] = acc[r] + t;
                }
            }
        }
        // out0 = (f = 0) + (f = 1): the f = 1 threads pass their sums through LDS
        cx<T>* ex = bufA;
        lds_barrier();
        if (f == 1 && j < NRL) {
#pragma unroll
            for (int r = 0; r < RL; ++r) ex[lds_addr<false, SZ, 0>(g, j + r * NRL)] = acc[r];
        }
        lds_barrier();
        if (row < nrows && j < NRL && f != 1) {
            const int q = f == 0 ? 0 : f - 1;
            cx<TS>* o = reinterpret_cast<cx<TS>*>(a.out) + ((size_t)grp * 3 + q) * a.ncell + cell0 + (uint32_t)g * N + j;
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                cx<T> val = acc[r];
                if (f == 0) val = val + ex[lds_addr<false, SZ, 0>(g, j + r * NRL)];
                o[r * NRL] = cx_as<TS>(val);
            }
        }
        lds_barrier();
        par = 0;
    }
}

// ---------------------------------------------------------------------------
// Source pass over source groups, one wave per row (64 / TPL rows per wave).
//
// Plans in the source-group form transform one representative atom per group in pass 0
// with the identity coefficient: [A, B, C] = F0[E - i k0 U_d0, U_d1, U_d2].  Lane (g, j)
// of a wave computes the source of cells j + r N / R0 of its row g, which are exactly the
// inputs of its stage-0 butterfly, so the source never passes through LDS.  The four
// fields then run one after another through the row's own LDS line.  Every hand-off is
// between lanes of one wave, and the LDS serves a wave's instructions in program order,
// so a compiler-level fence (wave_sync) orders them and no workgroup barrier is needed.
// A, B and C are stored from registers.  With one-hot phases both groups of a pair come
// from one read of the cells (see the one-hot branch); fractional phases take one group
// per read.  The twiddles of stage s sit in LDS as entry (r - 1) NS + k, so the lanes of a
// wave read consecutive entries.
constexpr int kSrcWaves = 4;  // waves per block of the one-wave-per-row source pass

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stage S of one line held in v (the same arithmetic as stage_compute).
template <typename T, class SZ, int S>
__device__ __forceinline__ void stage_compute_tw(cx<T> (&v)[SZ::RMAX], int j, const cx<T>* tw) {
    constexpr int R = SZ::rad(S), NS = SZ::ns(S);
    constexpr int OFF = StageTw<SZ>::off(S);
    if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[OFF + (r - 1) * NS + k]);
    }
    dft<R>(v);
}

// All stages of one line: the stage-0 inputs x[j + r N/R0] are in v on entry, the outputs
// X[j + r N/RL] on return.  act: the lane owns a line (inactive lanes touch no memory).
// P > 1: the last stage is split over P lanes per butterfly (last_stage_split) and the lane's
// outputs are those of LineOut<SZ, P>.
template <typename T, class SZ, int P = 1>
__device__ __forceinline__ void wave_line_transform(cx<T>* line, cx<T> (&v)[SZ::RMAX], int j, bool act,
                                                    const cx<T>* tw) {
    constexpr int NST = SZ::NST;
    static_assert(P == 1 || NST == 2, "split last stages are built for two-stage lines");
    wave_sync();  // the previous line's reads come before this line's writes
    const bool a0 = act && j < SZ::nr(0);
    if (a0) stage_compute_tw<T, SZ, 0>(v, j, tw);
    if constexpr (NST > 1) {
        if (a0) {
            constexpr int R = SZ::rad(0);
#pragma unroll
            for (int r = 0; r < R; ++r) line[lds_addr<false, SZ, 0>(0, j * R + r)] = v[r];
        }
        wave_sync();
        if constexpr (P > 1) {
            if (act && j < P * SZ::nr(1)) last_stage_split<T, SZ, P>(line, v, j % SZ::nr(1), j / SZ::nr(1), tw);
            return;
        }
        const bool a1 = act && j < SZ::nr(1);
        if (a1) {
            constexpr int R = SZ::rad(1), NR = SZ::nr(1), NS = SZ::ns(1);
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = line[lds_addr<false, SZ, 0>(0, j + r * NR)];
            stage_compute_tw<T, SZ, 1>(v, j, tw);
            if constexpr (NST > 2) {
                const int o = (j / NS) * NS * R + (j % NS);
#pragma unroll
                for (int r = 0; r < R; ++r) line[lds_addr<false, SZ, 0>(0, o + r * NS)] = v[r];
            }
        }
        if constexpr (NST > 2) {
            wave_sync();
            const bool a2 = act && j < SZ::nr(2);
            if (a2) {
                constexpr int R = SZ::rad(2), NR = SZ::nr(2), NS = SZ::ns(2);
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = line[lds_addr<false, SZ, 0>(0, j + r * NR)];
                stage_compute_tw<T, SZ, 2>(v, j, tw);
                if constexpr (NST > 3) {
                    const int o = (j / NS) * NS * R + (j % NS);
#pragma unroll
                    for (int r = 0; r < R; ++r) line[lds_addr<false, SZ, 0>(0, o + r * NS)] = v[r];
                }
            }
            if constexpr (NST > 3) {
                wave_sync();
                const bool a3 = act && j < SZ::nr(3);
                if (a3) {
                    constexpr int R = SZ::rad(3), NR = SZ::nr(3);
#pragma unroll
                    for (int r = 0; r < R; ++r) v[r] = line[lds_addr<false, SZ, 0>(0, j + r * NR)];
                    stage_compute_tw<T, SZ, 3>(v, j, tw);
                }
            }
        }
    }
}

// Three waves per SIMD: the register budget (168) that holds a row's eight cells of raw
// inputs and the four fields' stage-0 values with at most a couple of spilled values
// (measured faster than two waves without spills or four with more spills).  Lengths with
// a radix-11, 13 or 17 stage, whose DFTs need more registers, run two waves per SIMD instead
// of spilling; FP64 runs one, whose 512 registers (AGPRs included) hold a pair of groups'
// e and u_d without spills (at two waves it spilled 132 to 340 bytes a lane).
// IO32 (FP64 only): FP64 transforms on FP32 P, weights and outputs; the source is formed as in
// the FP32 kernel (FP64 phase reduction, FP32 values, held in FP32 registers) and each field
// is widened just before its transform, which keeps two waves per SIMD.
// P > 1: the last stage (prime radix) is split over P lanes per butterfly (LineOut).
template <typename T, class SZ, int ROWS, bool IO32 = false, int P = 1>
__global__ __launch_bounds__(64 * kSrcWaves) __attribute__((amdgpu_waves_per_eu(
    sizeof(T) == 8 ? (IO32 ? 2 : 1) : (SZ::RMAX >= 11 ? 2 : 3)))) void fast_source_wave(PassArgs<T> a) {
    constexpr bool F32 = std::is_same<T, float>::value;
    constexpr bool SRC32 = F32 || IO32;  // the source formed in FP32 (FP64 phase reduction)
    using TS = io_t<T, IO32>;
    using TT = std::conditional_t<SRC32, float, T>;  // atom-table entries
    constexpr int N = SZ::N, TPL = SZ::TPL, R0 = SZ::rad(0), NR0 = SZ::nr(0);
    constexpr int NTHR = 64 * kSrcWaves;
    constexpr int NTW = StageTw<SZ>::total();
    static_assert(ROWS >= 1 && ROWS * TPL <= 64, "a wave holds whole rows");
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lines = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* tw = lines + kSrcWaves * ROWS * SZ::PITCH;
    // the wavenumbers along the row, read from LDS where the outputs are formed: a global load
    // there, after the other fields' stores, would wait for those stores
    T* kl = reinterpret_cast<T*>(tw + NTW);
    TT* tabl = reinterpret_cast<TT*>(kl + N);
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int g = lane / TPL, j = lane - g * TPL;
    const bool inrow = g < ROWS;
    const uint32_t nrows = a.n1 * a.n2;
    const uint32_t row = (blockIdx.x * (uint32_t)kSrcWaves + (uint32_t)wave) * ROWS + (uint32_t)g;
    const bool live = inrow && row < nrows;
    fill_stage_tw<T, SZ, 0>(tw, a.W, tid, NTHR);
    fill_stage_tw<T, SZ, 1>(tw, a.W, tid, NTHR);
    fill_stage_tw<T, SZ, 2>(tw, a.W, tid, NTHR);
    fill_stage_tw<T, SZ, 3>(tw, a.W, tid, NTHR);
    for (int m = tid; m < N; m += NTHR) kl[m] = a.k[m];
    // FP32 source: the atom table in FP32 with the su rows in memory-dim order (entry
    // 2 + 3 m + k <- 2 + 3 ax_m + k); FP64 keeps the table's own order (cell_source picks the axes)
    for (int m = tid; m < a.n_atom_tab; m += NTHR) {
        const int e = m % kAtomTab;
        int src = m;
        if (SRC32 && e >= 2 && e < 11) {
            const int dm = (e - 2) / 3;
            const int ax = dm == 0 ? a.ax0 : (dm == 1 ? a.ax1 : a.ax2);
            src = m - e + 2 + 3 * ax + (e - 2) % 3;
        }
        tabl[m] = (TT)a.atomtab[src];
    }
    lds_barrier();
    cx<T>* line = lines + (wave * ROWS + (inrow ? g : 0)) * SZ::PITCH;
    const uint32_t cell0 = row * (uint32_t)N;

    const bool has0 = live && j < NR0;  // the lane owns stage-0 inputs
    // the lane's outputs: slot i of a line transform is point out_at(i) of the row (< 0: none)
    using LO = LineOut<SZ, P>;
    constexpr int NOS = LO::SLOTS;
    const bool has_out = live && LO::owns(j);
    int out_at[NOS];
#pragma unroll
    for (int i = 0; i < NOS; ++i) out_at[i] = LO::index(j, i);
    // fractional phases: transform group grp's four fields held in v (F[E] waits in
    // registers for F[U_d0]) and store A, B, C
    auto emit = [&](cx<T>(&v)[4][SZ::RMAX], int grp) {
        cx<TS>* oA = reinterpret_cast<cx<TS>*>(a.out) + (size_t)grp * 3 * a.ncell + (size_t)cell0;
        cx<TS>* oB = oA + a.ncell;
        cx<TS>* oC = oB + a.ncell;
        wave_line_transform<T, SZ, P>(line, v[0], j, inrow, tw);
        wave_line_transform<T, SZ, P>(line, v[1], j, inrow, tw);
        if (has_out) {
#pragma unroll
            for (int i = 0; i < NOS; ++i)
                if (P == 1 || out_at[i] >= 0) oA[out_at[i]] = cx_as<TS>(v[0][i] + cscale(negi(v[1][i]), kl[out_at[i]]));
        }
        wave_line_transform<T, SZ, P>(line, v[2], j, inrow, tw);
        if (has_out) {
#pragma unroll
            for (int i = 0; i < NOS; ++i)
                if (P == 1 || out_at[i] >= 0) oB[out_at[i]] = cx_as<TS>(v[2][i]);
        }
        wave_line_transform<T, SZ, P>(line, v[3], j, inrow, tw);
        if (has_out) {
#pragma unroll
            for (int i = 0; i < NOS; ++i)
                if (P == 1 || out_at[i] >= 0) oC[out_at[i]] = cx_as<TS>(v[3][i]);
        }
    };
    // one-hot phases: a group's fields are e and e u_d; each field e u_d is formed just
    // before its transform (in T from the held e and u, FP32 for an FP32 source), and
    // F[e u_0] waits for F[e], which is transformed last
    auto emit_eu = [&](auto& e, auto& u, int grp, uint32_t lane_cell) {
        // lane_cell - j is the row's first cell
        cx<TS>* oA = reinterpret_cast<cx<TS>*>(a.out) + (size_t)grp * 3 * a.ncell + (size_t)(lane_cell - (uint32_t)j);
        cx<TS>* oB = oA + a.ncell;
        cx<TS>* oC = oB + a.ncell;
        cx<T> w[SZ::RMAX], f0[NOS];
#pragma unroll
        for (int r = 0; r < R0; ++r) w[r] = cscale(cx_as<T>(e[r]), (T)u[0][r]);
        wave_line_transform<T, SZ, P>(line, w, j, inrow, tw);
#pragma unroll
        for (int i = 0; i < NOS; ++i) f0[i] = w[i];
#pragma unroll
        for (int r = 0; r < R0; ++r) w[r] = cscale(cx_as<T>(e[r]), (T)u[1][r]);
        wave_line_transform<T, SZ, P>(line, w, j, inrow, tw);
        if (has_out) {
#pragma unroll
            for (int i = 0; i < NOS; ++i)
                if (P == 1 || out_at[i] >= 0) oB[out_at[i]] = cx_as<TS>(w[i]);
        }
#pragma unroll
        for (int r = 0; r < R0; ++r) w[r] = cscale(cx_as<T>(e[r]), (T)u[2][r]);
        wave_line_transform<T, SZ, P>(line, w, j, inrow, tw);
        if (has_out) {
#pragma unroll
            for (int i = 0; i < NOS; ++i)
                if (P == 1 || out_at[i] >= 0) oC[out_at[i]] = cx_as<TS>(w[i]);
        }
#pragma unroll
        for (int r = 0; r < R0; ++r) w[r] = cx_as<T>(e[r]);
        wave_line_transform<T, SZ, P>(line, w, j, inrow, tw);
        if (has_out) {
#pragma unroll
            for (int i = 0; i < NOS; ++i)
                if (P == 1 || out_at[i] >= 0) oA[out_at[i]] = cx_as<TS>(w[i] + cscale(negi(f0[i]), kl[out_at[i]]));
        }
    };
    if constexpr (!SRC32) {
        // FP64: the phase of each atom formed whole in FP64; one-hot phases keep e and u_d
        // (5 values a cell) of a pair of groups per read of the cells, fractional ones the
        // four fields of one group
        for (int grp = 0; grp < a.ngroups; ++grp) {
            const int src = uload(a.chunk_src, uload(a.group_first, grp));
            // at most kAhead cells' loads in flight: cell r + kAhead is requested once cell r
            // has been formed (the scheduling barriers keep the compiler from hoisting them all)
            constexpr int kAhead = R0 < 2 ? R0 : 2;
            CellInputs<T> ci[R0];
#pragma unroll
            for (int r = 0; r < kAhead; ++r) cell_load<T, IO32>(a, cell0 + (uint32_t)(j + r * NR0), has0, ci[r]);
            if (!a.fraction) {
                // a pair of groups from one read of the cells
                const bool pair = grp + 1 < a.ngroups;
                const int srcB = pair ? uload(a.chunk_src, uload(a.group_first, grp + 1)) : src;
                cx<T> eA[SZ::RMAX], eB[SZ::RMAX];
                T uA[3][R0], uB[3][R0];
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    __builtin_amdgcn_sched_barrier(0);
                    if (r + kAhead < R0)
                        cell_load<T, IO32>(a, cell0 + (uint32_t)(j + (r + kAhead) * NR0), has0, ci[r + kAhead]);
                    cell_source_eu(a, tabl, ci[r], src, eA[r], uA[0][r], uA[1][r], uA[2][r]);
                    if (pair) cell_source_eu(a, tabl, ci[r], srcB, eB[r], uB[0][r], uB[1][r], uB[2][r]);
                }
                __builtin_amdgcn_sched_barrier(0);
                emit_eu(eA, uA, grp, cell0 + (uint32_t)j);
                if (pair) emit_eu(eB, uB, grp + 1, cell0 + (uint32_t)j);
                grp += pair ? 1 : 0;  // the pair's second group is done as well
            } else {
                cx<T> v[4][SZ::RMAX];
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    __builtin_amdgcn_sched_barrier(0);
                    if (r + kAhead < R0)
                        cell_load<T, IO32>(a, cell0 + (uint32_t)(j + (r + kAhead) * NR0), has0, ci[r + kAhead]);
                    cell_source<T, IO32>(a, tabl, ci[r], cell0 + (uint32_t)(j + r * NR0), src, v[0][r], v[1][r], v[2][r],
                                         v[3][r]);
                }
                __builtin_amdgcn_sched_barrier(0);
                emit(v, grp);
            }
        }
        return;
    }
    // The cells are formed one at a time after their loads; the scheduling barriers keep each
    // cell's atom-table reads next to its use.
    else if (!a.fraction) {
        // One-hot phases: a group's fields are e and e u_d, so a pair of groups is formed
        // from one read of the cells and only e and u_d (5 values a cell) are kept.  Each
        // field e u_d is formed just before its transform; F[e u_0] waits for F[e], which is
        // transformed last, in place.
        const PassArgs<float> af = source_view(a);
        for (int grp = 0; grp < a.ngroups; grp += 2) {
            const bool pair = grp + 1 < a.ngroups;
            const int srcA = uload(a.chunk_src, uload(a.group_first, grp));
            const int srcB = pair ? uload(a.chunk_src, uload(a.group_first, grp + 1)) : srcA;
            // the lane's first cell is re-formed in every iteration (the empty asm hides its
            // value), so only this 32-bit offset, not the 64-bit cell and output pointers
            // made from it, is held across the transforms
            uint32_t lane_cell = cell0 + (uint32_t)j;
            asm volatile("" : "+v"(lane_cell));
            const CellPtrsF cp = cell_ptrs_f(af, (size_t)lane_cell);
            cx<float> eA[SZ::RMAX], eB[SZ::RMAX];
            float uA[3][R0], uB[3][R0];
            // at most kAhead cells' loads in flight: cell r + kAhead is requested once cell r
            // has been formed
            constexpr int kAhead = R0 < 4 ? R0 : 4;
            RawCellF raw[R0];
#pragma unroll
            for (int r = 0; r < kAhead; ++r) raw_load_f(cp, r * NR0, has0, raw[r]);
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                __builtin_amdgcn_sched_barrier(0);
                CellInputsF ci;
                cell_prepare_f(af, raw[r], ci);
                if (r + kAhead < R0) raw_load_f(cp, (r + kAhead) * NR0, has0, raw[r + kAhead]);
                cell_source_1h(af, tabl, ci, srcA, eA[r], uA[0][r], uA[1][r], uA[2][r]);
                if (pair) cell_source_1h(af, tabl, ci, srcB, eB[r], uB[0][r], uB[1][r], uB[2][r]);
            }
            __builtin_amdgcn_sched_barrier(0);
            emit_eu(eA, uA, grp, lane_cell);
            if (pair) emit_eu(eB, uB, grp + 1, lane_cell);
        }
    } else {
        // fractional phases: one group per read of the cells
        const PassArgs<float> af = source_view(a);
        const CellPtrsF cp = cell_ptrs_f(af, (size_t)cell0 + (size_t)j);
        for (int grp = 0; grp < a.ngroups; ++grp) {
            const int src = uload(a.chunk_src, uload(a.group_first, grp));
            cx<T> v[4][SZ::RMAX];
            RawCellF raw[R0];
#pragma unroll
            for (int r = 0; r < R0; ++r) raw_load_f(cp, r * NR0, has0, raw[r]);
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                __builtin_amdgcn_sched_barrier(0);
                CellInputsF ci;
                cell_prepare_f(af, raw[r], ci);
                cx<float> s0, s1, s2, s3;
                cell_source_m(af, tabl, ci, cell0 + (uint32_t)(j + r * NR0), src, s0, s1, s2, s3);
                v[0][r] = cx_as<T>(s0);
                v[1][r] = cx_as<T>(s1);
                v[2][r] = cx_as<T>(s2);
                v[3][r] = cx_as<T>(s3);
            }
            __builtin_amdgcn_sched_barrier(0);
            emit(v, grp);
        }
    }
}

// ---------------------------------------------------------------------------
// Strided pass: block = G consecutive i0 (columns) x one outer index; thread
// (column g = tid % G, j = tid / G).  INTENSITY: write fftshifted max(|A|^2, 1e-50).
// Rows follow the split addressing of PassArgs (plain passes: nq = 1, q = 0, unit steps).
// PF: the next input is prefetched into registers while the current one is transformed;
// without it the kernel needs fewer registers and more blocks fit on a CU.
// IO32 (FP64 only): FP64 arithmetic on complex FP32 inputs and outputs.
// ONEBUF: one LDS stage buffer (transform's ONEBUF), for more blocks per CU.
// LDSIN (FP64, one LDS buffer): the next input is staged in LDS by sixteen-byte
// global_load_lds (as fast_mix_held_lds does) instead of registers; the item's coefficients
// and wavenumbers are loaded into registers before the next input's loads are issued, so
// that waiting for them does not wait for those.  n0 and the strides even.
template <typename T, class SZ, int G, int NIN, bool INTENSITY, bool SPLIT, bool PF = true, bool IO32 = false,
          bool ONEBUF = false, bool LDSIN = false>
__global__ __launch_bounds__(G* SZ::TPL) void fast_strided(PassArgs<T> a) {
    using TI = io_t<T, IO32>;
    static_assert(!LDSIN || (std::is_same<T, double>::value && ONEBUF && !PF), "staged inputs: FP64, one buffer");
    constexpr int EPL = 16 / (int)sizeof(cx<TI>);          // elements a staged lane loads
    constexpr int LPR = G / EPL > 0 ? G / EPL : 1;          // lanes per staged row
    constexpr int RPI = 64 / LPR;                           // staged rows per wave load
    constexpr int NLD = (SZ::N + RPI - 1) / RPI;            // wave loads per input
    constexpr int NWL = G * SZ::TPL / 64;                   // whole waves, which issue them
    static_assert(!LDSIN || (G % EPL == 0 && NWL >= 1), "staged rows of whole lanes, loaded by whole waves");
    __shared__ __align__(16) cx<TI> stg[LDSIN ? NLD * RPI * G : 1];
    // the NIN inputs of a chunk are transformed one after another (one line per thread)
    constexpr int N = SZ::N, R0 = SZ::rad(0), NR0 = SZ::nr(0), LP = SZ::strided_lp(G);
    constexpr int NOUT = NIN > 1 ? NIN - 1 : 1;
    constexpr int TILE = SZ::strided_tile(G);
    constexpr bool ALT = !ONEBUF && (chunk_writes<SZ, false>() & 1) == 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* bufA = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* bufB = ONEBUF ? bufA : bufA + TILE;
    cx<T>* Wl = bufB + TILE;
    const int tid = threadIdx.x;
    const int g = tid % G, j = tid / G;
    // rows of 256 bytes or more gain nothing from the XCD order (measured: they lose a little)
    const uint32_t blk = G * sizeof(cx<T>) < 256 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t t0 = blk % a.ntiles0;
    const uint32_t outer_q = blk / a.ntiles0;
    // unsplit passes: q = 0 and unit steps, known at compile time
    const uint32_t outer = SPLIT ? fdiv(outer_q, a.divQ) : outer_q;
    const uint32_t q = SPLIT ? outer_q - outer * a.nq : 0u;
    const uint32_t in_step = SPLIT ? a.in_step : 1u, out_step = SPLIT ? a.out_step : 1u;
    const uint32_t c_step = SPLIT ? a.c_step : 1u;
    const uint32_t i0 = t0 * G + g;
    const bool col_ok = i0 < a.n0;
    const uint32_t in0 = q * a.in_qmul, out0 = q * a.out_qmul, c0 = q * a.c_qmul;
    // j = jw + jl, where jw is the same on every lane of a wave (G divides 64): a row
    // address is a wave-uniform pointer (scalar registers) plus one 32-bit lane byte
    // offset (the host keeps these offsets below 4 GiB).  The idle columns of a partial
    // tile read the tile's last column: each column is transformed on its own, and
    // theirs are never stored.
    static_assert(64 % G == 0, "a wave holds whole rows of the tile");
    const uint32_t jw = (uint32_t)__builtin_amdgcn_readfirstlane(tid / 64) * (64 / G);
    const uint32_t jl = (uint32_t)(tid % 64) / G;
    const size_t tile0 = (size_t)t0 * G + (size_t)outer * a.OS;
    const uint32_t in_lane = ((col_ok ? i0 : a.n0 - 1) - t0 * G + jl * in_step * a.S) * (uint32_t)sizeof(cx<TI>);
    const uint32_t c_lane = jl * c_step;
    load_twiddles<T, SZ>(Wl, a.W, tid, G * SZ::TPL);
    lds_barrier();
    // inputs are consumed in the order (chunk, input)
    const int nitems = uload(a.group_first, a.ngroups) * NIN;
    // the prefetched line stays in the arrays' own type until it is used (a conversion at the
    // load would wait for it there)
    auto load_item = [&](int it, cx<TI>(&dst)[SZ::RMAX]) {
        const int c = it / NIN, f = it - (it / NIN) * NIN;
        const cx<TI>* in = reinterpret_cast<const cx<TI>*>(a.in) + ((size_t)uload(a.chunk_src, c) * a.nin + f) * a.ncell + tile0;
        if (j < NR0) {
#pragma unroll
            for (int r = 0; r < R0; ++r)
                dst[r] = load_at(in + (size_t)((in0 + (jw + r * NR0) * in_step) * a.S), in_lane);
        }
    };
    cx<TI> nxt[SZ::RMAX];
    if constexpr (PF) load_item(0, nxt);
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ln = tid & 63;
    // LDSIN: the wave loads of input it2 into stg (rows k RPI + ln / LPR, EPL columns from
    // EPL (ln % LPR)); the columns of a partial last tile beyond n0 are not loaded
    auto issue = [&](int it2) {
        const int c = it2 / NIN, f = it2 - (it2 / NIN) * NIN;
        const TI* in = reinterpret_cast<const TI*>(a.in) + 2 * (((size_t)uload(a.chunk_src, c) * a.nin + f) * a.ncell + tile0);
#pragma unroll
        for (int k0 = 0; k0 < NLD; k0 += NWL) {
            const int k = k0 + wv;
            const int row = k * RPI + ln / LPR;
            const uint32_t col = (uint32_t)(ln % LPR) * EPL;
            if (wv < NWL && k < NLD && row < SZ::N && t0 * G + col < a.n0)
                __builtin_amdgcn_global_load_lds(
                    (const void*)(in + 2 * ((size_t)((in0 + (uint32_t)row * in_step) * a.S) + (size_t)col)),
                    (__attribute__((address_space(3))) void*)(stg + k * RPI * G), 16, 0, 0);
        }
    };
    if constexpr (LDSIN) issue(0);
    int it = 0;
    int par = 0;
    for (int grp = 0; grp < a.ngroups; ++grp) {
        cx<T> acc[NOUT][SZ::RL];
#pragma unroll
        for (int o = 0; o < NOUT; ++o)
#pragma unroll
            for (int r = 0; r < SZ::RL; ++r) acc[o][r] = {(T)0, (T)0};
        for (int c = uload(a.group_first, grp); c < uload(a.group_first, grp + 1); ++c) {
            const int ct = uload(a.chunk_coef, c);
            const cx<T>* coef = a.coef + (ct >= 0 ? (size_t)ct * a.Nc : 0);
#pragma unroll
            for (int f = 0; f < NIN; ++f, ++it) {
                cx<T> v[1][SZ::RMAX];
                cx<T> cfv[LDSIN ? SZ::RL : 1];
                T kv[LDSIN ? SZ::RL : 1];
                if constexpr (LDSIN) {
                    __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // this wave's loads of input it
                    lds_barrier_nofence();                      // and every wave's
                    if (j < NR0) {
#pragma unroll
                        for (int r = 0; r < R0; ++r) v[0][r] = cx_as<T>(stg[(j + r * NR0) * G + g]);
                    }
                    if (j < SZ::NRL) {
                        const uint32_t coff = c_lane * (uint32_t)sizeof(cx<T>);
                        const uint32_t koff = c_lane * (uint32_t)sizeof(T);
#pragma unroll
                        for (int r = 0; r < SZ::RL; ++r) {
                            const uint32_t iu = c0 + (jw + r * SZ::NRL) * c_step;
                            cfv[r] = ct >= 0 ? load_at(coef + iu, coff) : cx<T>{(T)1, (T)0};
                            kv[r] = f == 1 ? load_at(a.k + iu, koff) : (T)0;
                        }
                    }
                    lds_barrier_nofence();  // every wave has read the staged tile
                    if (it + 1 < nitems) issue(it + 1);
                } else if constexpr (PF) {
#pragma unroll
                    for (int r = 0; r < SZ::RMAX; ++r) v[0][r] = cx_as<T>(nxt[r]);
                    if (it + 1 < nitems) load_item(it + 1, nxt);
                } else {
                    const cx<TI>* in =
                        reinterpret_cast<const cx<TI>*>(a.in) + ((size_t)uload(a.chunk_src, c) * a.nin + f) * a.ncell + tile0;
                    if (j < NR0) {
#pragma unroll
                        for (int r = 0; r < R0; ++r)
                            v[0][r] = cx_as<T>(load_at(in + (size_t)((in0 + (jw + r * NR0) * in_step) * a.S), in_lane));
                    }
                }
                transform<T, SZ, 1, true, LP, false, ONEBUF, LDSIN, !LDSIN>(par ? bufB : bufA, par ? bufA : bufB, v, j, g, 0, Wl);
                if (ALT) par ^= 1;
                if (j < SZ::NRL) {
                    const uint32_t coff = c_lane * (uint32_t)sizeof(cx<T>);
                    const uint32_t koff = c_lane * (uint32_t)sizeof(T);
#pragma unroll
                    for (int r = 0; r < SZ::RL; ++r) {
                        // coefficient index c0 + (j + r NRL) c_step, uniform part + lane part
                        const uint32_t iu = c0 + (jw + r * SZ::NRL) * c_step;
                        const cx<T> cf = LDSIN ? cfv[r] : (ct >= 0 ? load_at(coef + iu, coff) : cx<T>{(T)1, (T)0});
                        cx<T> t = cmul(cf, v[0][r]);
                        if (f == 1) t = cscale(negi(t), LDSIN ? kv[r] : load_at(a.k + iu, koff));
                        const int oq = f <= 1 ? 0 : f - 1;
                        acc[oq][r] = acc[oq][r] + t;
                    }
                }
            }
        }
        if (!col_ok || j >= SZ::NRL) continue;
        if constexpr (INTENSITY) {
            const uint32_t s0 = (i0 + a.n0 / 2) >= a.n0 ? i0 + a.n0 / 2 - a.n0 : i0 + a.n0 / 2;
            const uint32_t nA = a.dim == 2 ? a.n2 : a.n1;
            const uint32_t nB = a.dim == 2 ? a.n1 : a.n2;
            const uint32_t sB = (outer + nB / 2) >= nB ? outer + nB / 2 - nB : outer + nB / 2;
#pragma unroll
            for (int r = 0; r < SZ::RL; ++r) {
                const uint32_t idx = c0 + (uint32_t)(j + r * SZ::NRL) * c_step;
                const uint32_t sA = (idx + nA / 2) >= nA ? idx + nA / 2 - nA : idx + nA / 2;
                const uint32_t s1 = a.dim == 2 ? sB : sA;
                const uint32_t s2 = a.dim == 2 ? sA : sB;
                const double re = (double)acc[0][r].x, im = (double)acc[0][r].y;
                a.I[(size_t)s0 + (size_t)a.n0 * ((size_t)s1 + (size_t)a.n1 * s2)] = fmax(re * re + im * im, 1e-50);
            }
        } else {
            cx<TI>* o = reinterpret_cast<cx<TI>*>(a.out) + (size_t)grp * NOUT * a.ncell + tile0;
            const uint32_t out_lane = (g + jl * out_step * a.S) * (uint32_t)sizeof(cx<TI>);
#pragma unroll
            for (int qq = 0; qq < NOUT; ++qq)
#pragma unroll
                for (int r = 0; r < SZ::RL; ++r)
                    store_at(o + (size_t)qq * a.ncell + (size_t)((out0 + (jw + r * SZ::NRL) * out_step) * a.S), out_lane,
                             cx_as<TI>(acc[qq][r]));
        }
    }
}

// ---------------------------------------------------------------------------
// Pass 1 over source groups.  Atoms whose displacement coefficients su (and form-factor
// ratios across phases) are equal have proportional sources, fE_a = phi_a fE_S, so pass 0
// can transform one representative per group S and leave the atom coefficients for
// later: its outputs are [A_S, B_S, C_S] = F0 [E - i k0 U_0, U_1, U_2] of the group.
// With X_{Y,S}(k0) = sum over atoms a in class Y and group S of phi_a c_a(k0) (the
// pass-0 coefficients of the class form, a function of the column only), this pass
// forms for every z-class Z
//   out_{Z,0} = sum_S sum_{Y in Z} c_Y(k1) X_{Y,S}(k0) [F1 A_S - i k1 F1 B_S],
//   out_{Z,1} = sum_S sum_{Y in Z} c_Y(k1) X_{Y,S}(k0) F1 C_S,
// the same outputs as the class form.  Each group's three inputs are transformed once.
// IO32 (FP64 only): FP64 arithmetic on complex FP32 inputs and outputs.
// ONEBUF: one LDS stage buffer (transform's ONEBUF), for more blocks per CU.
// TABL: the block's wavenumbers, coefficient rows (a.ncoef of them) and mix factors are staged
// in LDS once (a.ncoef (N + a.nsrc G) values and N reals after the twiddles): read from global
// memory after the next input's prefetch is issued, each of them waited for that prefetch.
template <typename T, class SZ, int G, int NZ, bool SPLIT, bool IO32 = false, bool ONEBUF = false, bool TABL = false>
__global__ __launch_bounds__(G* SZ::TPL) void fast_strided_mix(PassArgs<T> a) {
    using TI = io_t<T, IO32>;
    constexpr int N = SZ::N, R0 = SZ::rad(0), NR0 = SZ::nr(0), LP = SZ::strided_lp(G), RL = SZ::RL, NRL = SZ::NRL;
    constexpr int TILE = SZ::strided_tile(G);
    constexpr bool ALT = !ONEBUF && (chunk_writes<SZ, false>() & 1) == 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* bufA = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* bufB = ONEBUF ? bufA : bufA + TILE;
    cx<T>* Wl = bufB + TILE;
    const int tid = threadIdx.x;
    const int g = tid % G, j = tid / G;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t t0 = blk % a.ntiles0;
    const uint32_t outer_q = blk / a.ntiles0;
    // unsplit passes: q = 0 and unit steps, known at compile time
    const uint32_t outer = SPLIT ? fdiv(outer_q, a.divQ) : outer_q;
    const uint32_t q = SPLIT ? outer_q - outer * a.nq : 0u;
    const uint32_t in_step = SPLIT ? a.in_step : 1u, out_step = SPLIT ? a.out_step : 1u;
    const uint32_t c_step = SPLIT ? a.c_step : 1u;
    const uint32_t in0 = q * a.in_qmul, out0 = q * a.out_qmul, c0 = q * a.c_qmul;
    const uint32_t i0 = t0 * G + g;
    const bool col_ok = i0 < a.n0;
    // row addresses: wave-uniform pointer + 32-bit lane offset, as in fast_strided
    static_assert(64 % G == 0, "a wave holds whole rows of the tile");
    const uint32_t jw = (uint32_t)__builtin_amdgcn_readfirstlane(tid / 64) * (64 / G);
    const uint32_t jl = (uint32_t)(tid % 64) / G;
    const size_t tile0 = (size_t)t0 * G + (size_t)outer * a.OS;
    const uint32_t in_lane = ((col_ok ? i0 : a.n0 - 1) - t0 * G + jl * in_step * a.S) * (uint32_t)sizeof(cx<TI>);
    const uint32_t c_lane = jl * c_step;
    load_twiddles<T, SZ>(Wl, a.W, tid, G * SZ::TPL);
    T* kl = reinterpret_cast<T*>(Wl + N);
    cx<T>* cyl = Wl + N + (N + 1) / 2;
    cx<T>* xcl = cyl + (size_t)a.ncoef * N;
    if constexpr (TABL) {
        // row idx of the line: wavenumber and coefficient index c0 + idx c_step
        for (int m = tid; m < N; m += G * SZ::TPL) kl[m] = a.k[c0 + (uint32_t)m * c_step];
        for (int m = tid; m < a.ncoef * N; m += G * SZ::TPL) {
            const int y = m / N, idx = m - y * N;
            cyl[m] = a.coef[(size_t)y * a.Nc + c0 + (uint32_t)idx * c_step];
        }
        for (int m = tid; m < a.ncoef * a.nsrc * G; m += G * SZ::TPL) {
            const int ys = m / G, c = m - ys * G;
            const uint32_t col = t0 * G + (uint32_t)c;
            xcl[m] = col < a.n0 ? a.mixc[(size_t)ys * a.n0 + col] : cx<T>{(T)0, (T)0};
        }
    }
    lds_barrier();
    constexpr bool PF = true;
    const int nitems = 3 * a.nsrc;
    auto load_item = [&](int it, cx<TI>(&dst)[SZ::RMAX]) {
        const cx<TI>* in = reinterpret_cast<const cx<TI>*>(a.in) + (size_t)it * a.ncell + tile0;
        if (j < NR0) {
#pragma unroll
            for (int r = 0; r < R0; ++r) dst[r] = load_at(in + (size_t)((in0 + (jw + r * NR0) * in_step) * a.S), in_lane);
        }
    };
    cx<T> acc[NZ][2][RL];
#pragma unroll
    for (int z = 0; z < NZ; ++z)
#pragma unroll
        for (int r = 0; r < RL; ++r) acc[z][0][r] = acc[z][1][r] = {(T)0, (T)0};
    cx<TI> nxt[SZ::RMAX];
    if constexpr (PF) load_item(0, nxt);
    int it = 0;
    int par = 0;
    for (int s = 0; s < a.nsrc; ++s) {
        cx<T> tA[RL], tC[RL];
#pragma unroll
        for (int f = 0; f < 3; ++f, ++it) {
            cx<T> v[1][SZ::RMAX];
            if constexpr (PF) {
#pragma unroll
                for (int r = 0; r < SZ::RMAX; ++r) v[0][r] = cx_as<T>(nxt[r]);
                if (it + 1 < nitems) load_item(it + 1, nxt);
            } else {
                cx<TI> cur[SZ::RMAX];
                load_item(it, cur);
#pragma unroll
                for (int r = 0; r < SZ::RMAX; ++r) v[0][r] = cx_as<T>(cur[r]);
            }
            transform<T, SZ, 1, true, LP, false, ONEBUF>(par ? bufB : bufA, par ? bufA : bufB, v, j, g, 0, Wl);
            if (ALT) par ^= 1;
            if (j < NRL) {
                const uint32_t koff = c_lane * (uint32_t)sizeof(T);
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    if (f == 0) tA[r] = v[0][r];
                    if (f == 1)
                        tA[r] = tA[r] + cscale(negi(v[0][r]), TABL ? kl[j + r * NRL] : load_at(a.k + c0 + (jw + r * NRL) * c_step, koff));
                    if (f == 2) tC[r] = v[0][r];
                }
            }
        }
        // the pair lists are walked by the whole wave (their entries stay in scalar
        // registers); only the lanes of stored points accumulate
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            const int p1 = uload(a.pair_first, s * NZ + z + 1);
            for (int p = uload(a.pair_first, s * NZ + z); p < p1; ++p) {
                const int y = uload(a.pair_y, p);
                const cx<T>* xrow = a.mixc + (size_t)(y * a.nsrc + s) * a.n0 + t0 * G;
                const cx<T>* cy = a.coef + (size_t)y * a.Nc + c0;
                if (j < NRL && col_ok) {
                    const cx<T> xc = TABL ? xcl[(y * a.nsrc + s) * G + g] : load_at(xrow, g * (uint32_t)sizeof(cx<T>));
                    const uint32_t coff = c_lane * (uint32_t)sizeof(cx<T>);
#pragma unroll
                    for (int r = 0; r < RL; ++r) {
                        const cx<T> cf = cmul(TABL ? cyl[y * N + j + r * NRL] : load_at(cy + (jw + r * NRL) * c_step, coff), xc);
                        acc[z][0][r] = acc[z][0][r] + cmul(cf, tA[r]);
                        acc[z][1][r] = acc[z][1][r] + cmul(cf, tC[r]);
                    }
                }
            }
        }
    }
    if (!col_ok || j >= NRL) return;
    const uint32_t out_lane = (g + jl * out_step * a.S) * (uint32_t)sizeof(cx<TI>);
#pragma unroll
    for (int z = 0; z < NZ; ++z)
#pragma unroll
        for (int oq = 0; oq < 2; ++oq) {
            cx<TI>* o = reinterpret_cast<cx<TI>*>(a.out) + (size_t)(2 * z + oq) * a.ncell + tile0;
#pragma unroll
            for (int r = 0; r < RL; ++r)
                store_at(o + (size_t)((out0 + (jw + r * NRL) * out_step) * a.S), out_lane, cx_as<TI>(acc[z][oq][r]));
        }
}

// ---------------------------------------------------------------------------
// Pass 1 over NS source groups (NS known at compile time), forming the same outputs as
// fast_strided_mix with fewer values held per thread.  With the column's mix factors
// M_{Z,S}(k0, k1) = sum over y-classes Y in Z of c_Y(k1) X_{Y,S}(k0),
//   out_{Z,0} = sum_S M_{Z,S} T_S,  T_S = F1 A_S - i k1 F1 B_S,
//   out_{Z,1} = sum_S M_{Z,S} F1 C_S.
// The inputs are taken in the order A_0 B_0 A_1 B_1 ... then C_0 C_1 ...: once every
// group's T_S is held, the first outputs of all z-classes are formed and stored, and the
// held values are reused for the F1 C_S.  A thread holds NS RL group values (the last
// transform of a round runs beside NS - 1 of them), where fast_strided_mix holds NZ 2 RL
// sums and 2 RL group values; the prefetched line stays in the arrays' own type, and the
// wavenumbers, the coefficient tables and the tile's mix factors are read from LDS (vector
// loads of them in the mix waited for the prefetch too: the mix took 0.4 of 0.87 ms).  The
// launch adds a.ncoef (N + NS G) values to the kernel's LDS.  PF: the next input is
// prefetched into registers during the current transform; without it the kernel holds fewer
// registers and is built for at least three waves per SIMD.
template <typename T, class SZ, int G, int NZ, int NS, bool IO32 = false, bool ONEBUF = false, bool PF = true>
__global__ __launch_bounds__(G* SZ::TPL) __attribute__((amdgpu_waves_per_eu(PF ? 1 : 3))) void fast_mix_held(
    PassArgs<T> a) {
    using TI = io_t<T, IO32>;
    constexpr int N = SZ::N, R0 = SZ::rad(0), NR0 = SZ::nr(0), LP = SZ::strided_lp(G), RL = SZ::RL, NRL = SZ::NRL;
    constexpr int TILE = SZ::strided_tile(G);
    constexpr bool ALT = !ONEBUF && (chunk_writes<SZ, false>() & 1) == 1;
    constexpr int NITEMS = 3 * NS;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* bufA = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* bufB = ONEBUF ? bufA : bufA + TILE;
    cx<T>* Wl = bufB + TILE;
    T* kl = reinterpret_cast<T*>(Wl + N);  // the wavenumbers along the pass, read from LDS
    // the pass's coefficient tables (a.ncoef x N) and the tile's mix factors (a.ncoef x NS x G),
    // read from LDS by the mix
    cx<T>* cyl = Wl + N + (N + 1) / 2;  // after the wavenumbers, aligned to a complex value
    cx<T>* xcl = cyl + (size_t)a.ncoef * N;
    const int tid = threadIdx.x;
    const int g = tid % G, j = tid / G;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t t0 = blk % a.ntiles0;
    const uint32_t outer = blk / a.ntiles0;
    const uint32_t i0 = t0 * G + g;
    const bool col_ok = i0 < a.n0;
    // row addresses: wave-uniform pointer + 32-bit lane offset, as in fast_strided
    static_assert(64 % G == 0, "a wave holds whole rows of the tile");
    const uint32_t jw = (uint32_t)__builtin_amdgcn_readfirstlane(tid / 64) * (64 / G);
    const uint32_t jl = (uint32_t)(tid % 64) / G;
    const size_t tile0 = (size_t)t0 * G + (size_t)outer * a.OS;
    const uint32_t in_lane = ((col_ok ? i0 : a.n0 - 1) - t0 * G + jl * a.S) * (uint32_t)sizeof(cx<TI>);
    const uint32_t out_lane = (g + jl * a.S) * (uint32_t)sizeof(cx<TI>);
    load_twiddles<T, SZ>(Wl, a.W, tid, G * SZ::TPL);
    for (int m = tid; m < N; m += G * SZ::TPL) kl[m] = a.k[m];
    for (int m = tid; m < a.ncoef * N; m += G * SZ::TPL) cyl[m] = a.coef[m];
    for (int m = tid; m < a.ncoef * NS * G; m += G * SZ::TPL) {
        const int ys = m / G, c = m - ys * G;
        const uint32_t col = t0 * G + (uint32_t)c;
        xcl[m] = col < a.n0 ? a.mixc[(size_t)ys * a.n0 + col] : cx<T>{(T)0, (T)0};
    }
    lds_barrier();
    // item it -> input field: A_s, B_s for s = it / 2 while it < 2 NS, then C_{it - 2 NS}
    auto load_item = [&](int it, cx<TI>(&dst)[SZ::RMAX]) {
        const int fld = it < 2 * NS ? (it >> 1) * 3 + (it & 1) : (it - 2 * NS) * 3 + 2;
        const cx<TI>* in = reinterpret_cast<const cx<TI>*>(a.in) + (size_t)fld * a.ncell + tile0;
        if (j < NR0) {
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                // the row offsets stay uniform values even where the compiler keeps them in
                // vector registers (no scalar registers left), so the loads keep one scalar
                // resource instead of a loop over the lanes' resources
                const uint32_t off = __builtin_amdgcn_readfirstlane((jw + r * NR0) * a.S);
                dst[r] = load_at(in + (size_t)off, in_lane);
            }
        }
    };
    // out_{Z,oq} = sum_S M_{Z,S} held[S] for every z-class, stored at once.  The pair lists
    // are walked by the whole wave (their entries stay in scalar registers).
    auto mix_store = [&](int oq, const cx<T>(&held)[NS][RL]) {
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            cx<T> acc[RL];
#pragma unroll
            for (int r = 0; r < RL; ++r) acc[r] = {(T)0, (T)0};
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int p1 = uload(a.pair_first, s * NZ + z + 1);
                for (int p = uload(a.pair_first, s * NZ + z); p < p1; ++p) {
                    const int y = uload(a.pair_y, p);
                    const cx<T>* cy = cyl + y * N;
                    if (j < NRL && col_ok) {
                        const cx<T> xc = xcl[(y * NS + s) * G + g];
#pragma unroll
                        for (int r = 0; r < RL; ++r) acc[r] = acc[r] + cmul(cmul(cy[j + r * NRL], xc), held[s][r]);
                    }
                }
            }
            if (j < NRL && col_ok) {
                cx<TI>* o = reinterpret_cast<cx<TI>*>(a.out) + (size_t)(2 * z + oq) * a.ncell + tile0;
#pragma unroll
                for (int r = 0; r < RL; ++r) store_at(o + (size_t)((jw + r * NRL) * a.S), out_lane, cx_as<TI>(acc[r]));
            }
        }
    };
    cx<T> held[NS][RL];
    cx<TI> nxt[PF ? SZ::RMAX : 1];
    if constexpr (PF) load_item(0, nxt);
    int par = 0;
#pragma unroll 1
    for (int it = 0; it < NITEMS; ++it) {
        cx<T> v[1][SZ::RMAX];
        if constexpr (PF) {
#pragma unroll
            for (int r = 0; r < SZ::RMAX; ++r) v[0][r] = cx_as<T>(nxt[r]);
            if (it + 1 < NITEMS) load_item(it + 1, nxt);
        } else {
            cx<TI> cur[SZ::RMAX];
            load_item(it, cur);
#pragma unroll
            for (int r = 0; r < SZ::RMAX; ++r) v[0][r] = cx_as<T>(cur[r]);
        }
        transform<T, SZ, 1, true, LP, false, ONEBUF>(par ? bufB : bufA, par ? bufA : bufB, v, j, g, 0, Wl);
        if (ALT) par ^= 1;
        if (j < NRL) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (it == 2 * s || it == 2 * NS + s) {
#pragma unroll
                    for (int r = 0; r < RL; ++r) held[s][r] = v[0][r];
                } else if (it == 2 * s + 1) {
#pragma unroll
                    for (int r = 0; r < RL; ++r)
                        held[s][r] = held[s][r] + cscale(negi(v[0][r]), kl[j + r * NRL]);
                }
            }
        }
        if (it == 2 * NS - 1) mix_store(0, held);
        if (it == NITEMS - 1) mix_store(1, held);
    }
}

// ---------------------------------------------------------------------------
// Pass 1 over NS source groups as fast_mix_held (FP64 transforms on FP32 spectra, one LDS
// stage buffer), with the next input staged in LDS instead of registers: each wave loads its
// rows of the tile with global_load_lds_dwordx4 (sixteen bytes a lane, two columns of a row;
// a wave's load fills 1 KB of LDS in lane order, eight rows of sixteen columns), so the load
// of input it + 1 stays in flight across the transform of input it without holding registers.
// Only whole waves issue these loads (a wave's load writes all 64 lanes' slots; the partial
// last wave of a 13-row tile has 16 lanes).
// Every wave waits for its own loads (vmcnt(0)) before the barrier after which the tile is
// read, and the next loads are issued only after a second barrier, once every wave has read
// its stage-0 inputs.  n0 and the row stride must be even (16-byte aligned column pairs).
// T = float: the same in FP32 throughout.  IO32 false (FP64 plans): FP64 spectra, one value a
// staged lane (a row of G = 8 columns is 128 bytes).
template <class SZ, int G, int NZ, int NS, int WPE, typename T = double, bool IO32 = true>
__global__ __launch_bounds__(G* SZ::TPL) __attribute__((amdgpu_waves_per_eu(WPE))) void fast_mix_held_lds(
    PassArgs<T> a) {
    using TI = std::conditional_t<IO32, float, T>;
    constexpr int N = SZ::N, R0 = SZ::rad(0), NR0 = SZ::nr(0), LP = SZ::strided_lp(G), RL = SZ::RL, NRL = SZ::NRL;
    constexpr int TILE = SZ::strided_tile(G);
    constexpr int NITEMS = 3 * NS;
    constexpr int EPL = 16 / (int)sizeof(cx<TI>);   // values a staged lane loads
    static_assert(G % EPL == 0 && G / EPL <= 8, "a staged row is G / EPL lanes of sixteen bytes");
    constexpr int LPR = G / EPL;                    // lanes per staged row
    constexpr int RPI = 64 / LPR;                   // rows per wave load
    constexpr int NLD = (N + RPI - 1) / RPI;        // wave loads per input
    constexpr int NWAVES = G * SZ::TPL / 64;         // whole waves, which issue the loads
    static_assert(NWAVES >= 1, "a whole wave loads the tile");
    __shared__ __align__(16) cx<TI> stg[NLD * RPI * G];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* bufA = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* Wl = bufA + TILE;
    T* kl = reinterpret_cast<T*>(Wl + N);
    cx<T>* cyl = Wl + N + (N + 1) / 2;
    cx<T>* xcl = cyl + (size_t)a.ncoef * N;
    const int tid = threadIdx.x;
    const int g = tid % G, j = tid / G;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t t0 = blk % a.ntiles0;
    const uint32_t outer = blk / a.ntiles0;
    static_assert(64 % G == 0, "a wave holds whole rows of the tile");
    const uint32_t jw = (uint32_t)__builtin_amdgcn_readfirstlane(tid / 64) * (64 / G);
    const uint32_t jl = (uint32_t)(tid % 64) / G;
    const size_t tile0 = (size_t)t0 * G + (size_t)outer * a.OS;
    const uint32_t out_lane = (g + jl * a.S) * (uint32_t)sizeof(cx<TI>);
    const bool col_ok = t0 * G + (uint32_t)g < a.n0;
    load_twiddles<T, SZ>(Wl, a.W, tid, G * SZ::TPL);
    for (int m = tid; m < N; m += G * SZ::TPL) kl[m] = a.k[m];
    for (int m = tid; m < a.ncoef * N; m += G * SZ::TPL) cyl[m] = a.coef[m];
    for (int m = tid; m < a.ncoef * NS * G; m += G * SZ::TPL) {
        const int ys = m / G, c = m - ys * G;
        const uint32_t col = t0 * G + (uint32_t)c;
        xcl[m] = col < a.n0 ? a.mixc[(size_t)ys * a.n0 + col] : cx<T>{(T)0, (T)0};
    }
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ln = tid & 63;
    // input it -> field: A_s, B_s for s = it / 2 while it < 2 NS, then C_{it - 2 NS}
    auto issue = [&](int it, cx<TI>* buf) {
        const int fld = it < 2 * NS ? (it >> 1) * 3 + (it & 1) : (it - 2 * NS) * 3 + 2;
        const TI* in = reinterpret_cast<const TI*>(a.in) + 2 * ((size_t)fld * a.ncell + tile0);
#pragma unroll
        for (int k0 = 0; k0 < NLD; k0 += NWAVES) {
            const int k = k0 + wv;
            const int row = k * RPI + ln / LPR;
            // the columns of a partial last tile beyond n0 (even) are not loaded: their slots
            // keep what they held, their results are not stored
            const uint32_t c2 = (uint32_t)(ln % LPR) * EPL;
            if (wv < NWAVES && k < NLD && row < N && t0 * G + c2 < a.n0)
                __builtin_amdgcn_global_load_lds((const void*)(in + 2 * ((size_t)row * a.S + (size_t)c2)),
                                                 (__attribute__((address_space(3))) void*)(buf + k * RPI * G), 16, 0, 0);
        }
    };
    auto mix_store = [&](int oq, const cx<T>(&held)[NS][RL]) {
#pragma unroll
        for (int z = 0; z < NZ; ++z) {
            cx<T> acc[RL];
#pragma unroll
            for (int r = 0; r < RL; ++r) acc[r] = {(T)0, (T)0};
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const int p1 = uload(a.pair_first, s * NZ + z + 1);
                for (int p = uload(a.pair_first, s * NZ + z); p < p1; ++p) {
                    const int y = uload(a.pair_y, p);
                    const cx<T>* cy = cyl + y * N;
                    if (j < NRL) {
                        const cx<T> xc = xcl[(y * NS + s) * G + g];
#pragma unroll
                        for (int r = 0; r < RL; ++r) acc[r] = acc[r] + cmul(cmul(cy[j + r * NRL], xc), held[s][r]);
                    }
                }
            }
            if (j < NRL && col_ok) {
                cx<TI>* o = reinterpret_cast<cx<TI>*>(a.out) + (size_t)(2 * z + oq) * a.ncell + tile0;
#pragma unroll
                for (int r = 0; r < RL; ++r) store_at(o + (size_t)((jw + r * NRL) * a.S), out_lane, cx_as<TI>(acc[r]));
            }
        }
    };
    cx<T> held[NS][RL];
    // input it: wait for this wave's loads of it (vmcnt(0)), then a barrier for every wave's;
    // read the stage-0 inputs; a barrier before the buffer is loaded with input it + 1;
    // transform, update the held values, mix and store
    auto step = [&](int it, cx<TI>* buf) {
        __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
        lds_barrier_nofence();
        cx<T> v[1][SZ::RMAX];
        if (j < NR0) {
#pragma unroll
            for (int r = 0; r < R0; ++r) v[0][r] = cx_as<T>(buf[(j + r * NR0) * G + g]);
        }
        lds_barrier_nofence();  // every wave has read the staged tile
        if (it + 1 < NITEMS) issue(it + 1, buf);
        // the barrier after the staged tile was read also orders the previous input's last reads
        // of bufA before this input's first write
        transform<T, SZ, 1, true, LP, false, true, true, false>(bufA, bufA, v, j, g, 0, Wl);
        if (j < NRL) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                if (it == 2 * s || it == 2 * NS + s) {
#pragma unroll
                    for (int r = 0; r < RL; ++r) held[s][r] = v[0][r];
                } else if (it == 2 * s + 1) {
#pragma unroll
                    for (int r = 0; r < RL; ++r) held[s][r] = held[s][r] + cscale(negi(v[0][r]), kl[j + r * NRL]);
                }
            }
        }
        if (it == 2 * NS - 1) mix_store(0, held);
        if (it == NITEMS - 1) mix_store(1, held);
    };
    issue(0, stg);
#pragma unroll 1
    for (int it = 0; it < NITEMS; ++it) step(it, stg);
}

// ---------------------------------------------------------------------------
// First step of a split strided transform of length Nc = N1 N, in place on every
// field: for the line (column i0, q < N1, outer) the N points at rows q + N1 y2 are
// transformed (length N = SZ::N) and output k2 is multiplied by exp(-2 pi i q k2 / Nc)
// and written back to row q + N1 k2 (in_qmul = out_qmul = 1, in_step = out_step = N1).
// Each block reads all of a line before it writes it, and blocks own disjoint lines.
// IO32 (FP64 only): reads the FP32 fields at a.in and writes the FP64 results at the same
// places of a.out (a separate buffer), so the intermediate keeps FP64 precision.
// LDSIN (FP32 arrays in, G >= 2): the next field is staged in LDS by sixteen-byte
// global_load_lds (two columns of a row a lane, as fast_mix_held_lds does) instead of
// registers, and the block's twiddles W^{q k2} are read from LDS; n0 and the strides even.
template <typename T, class SZ, int G, bool IO32 = false, bool LDSIN = false>
__global__ __launch_bounds__(G* SZ::TPL) void fast_split(PassArgs<T> a) {
    using TI = io_t<T, IO32>;
    constexpr int N = SZ::N, R0 = SZ::rad(0), NR0 = SZ::nr(0), LP = SZ::strided_lp(G);
    constexpr int TILE = SZ::strided_tile(G);
    constexpr bool ALT = !LDSIN && (chunk_writes<SZ, false>() & 1) == 1;
    static_assert(!LDSIN || (std::is_same<TI, float>::value && G >= 2 && G <= 16), "staged FP32 rows");
    constexpr int LPR = G / 2 > 0 ? G / 2 : 1;   // lanes per staged row (sixteen bytes each)
    constexpr int RPI = 64 / LPR;                // staged rows per wave load
    constexpr int NLD = (N + RPI - 1) / RPI;     // wave loads per field
    constexpr int NWL = G * SZ::TPL / 64;        // whole waves, which issue them
    static_assert(!LDSIN || NWL >= 1, "a whole wave loads the tile");
    __shared__ __align__(16) cx<float> stg[LDSIN ? NLD * RPI * G : 1];
    __shared__ __align__(16) cx<T> wsl[LDSIN ? N : 1];
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* bufA = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* bufB = LDSIN ? bufA : bufA + TILE;
    cx<T>* Wl = bufA + (LDSIN ? 1 : 2) * TILE;
    const int tid = threadIdx.x;
    const int g = tid % G, j = tid / G;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t t0 = blk % a.ntiles0;
    const uint32_t outer_q = blk / a.ntiles0;
    const uint32_t outer = fdiv(outer_q, a.divQ);
    const uint32_t q = outer_q - outer * a.nq;
    const uint32_t i0 = t0 * G + g;
    const bool col_ok = i0 < a.n0;
    const uint32_t rstride = a.in_step * a.S;
    // row addresses: wave-uniform pointer + 32-bit lane offset, as in fast_strided
    static_assert(64 % G == 0, "a wave holds whole rows of the tile");
    const uint32_t jw = (uint32_t)__builtin_amdgcn_readfirstlane(tid / 64) * (64 / G);
    const uint32_t jl = (uint32_t)(tid % 64) / G;
    const size_t tile0 = (size_t)t0 * G + (size_t)outer * a.OS + (size_t)q * a.in_qmul * a.S;
    const uint32_t in_lane = ((col_ok ? i0 : a.n0 - 1) - t0 * G + jl * rstride) * (uint32_t)sizeof(cx<TI>);
    load_twiddles<T, SZ>(Wl, a.W, tid, G * SZ::TPL);
    if constexpr (LDSIN) {
        for (int m = tid; m < N; m += G * SZ::TPL) wsl[m] = a.Wsplit[q * (uint32_t)m];
    }
    lds_barrier();
    constexpr bool PF = !LDSIN;
    const cx<TI>* src = IO32 ? reinterpret_cast<const cx<TI>*>(a.in) : reinterpret_cast<const cx<TI>*>(a.out);
    auto load_field = [&](int f, cx<TI>(&dst)[SZ::RMAX]) {
        const cx<TI>* in = src + (size_t)f * a.ncell + tile0;
        if (j < NR0) {
#pragma unroll
            for (int r = 0; r < R0; ++r) dst[r] = load_at(in + (size_t)(jw + r * NR0) * rstride, in_lane);
        }
    };
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ln = tid & 63;
    // LDSIN: the wave loads of field f into stg (rows k RPI + ln / LPR, columns 2 (ln % LPR));
    // the columns of a partial last tile beyond n0 (even) are not loaded, nor stored
    auto issue = [&](int f) {
        const float* in = reinterpret_cast<const float*>(src) + 2 * ((size_t)f * a.ncell + tile0);
#pragma unroll
        for (int k0 = 0; k0 < NLD; k0 += NWL) {
            const int k = k0 + wv;
            const int row = k * RPI + ln / LPR;
            const uint32_t c2 = (uint32_t)(ln % LPR) * 2;
            if (wv < NWL && k < NLD && row < N && t0 * G + c2 < a.n0)
                __builtin_amdgcn_global_load_lds((const void*)(in + 2 * ((size_t)row * rstride + (size_t)c2)),
                                                 (__attribute__((address_space(3))) void*)(stg + k * RPI * G), 16, 0, 0);
        }
    };
    cx<TI> nxt[SZ::RMAX];
    if constexpr (PF) load_field(0, nxt);
    if constexpr (LDSIN) issue(0);
    int par = 0;
    for (int f = 0; f < a.nfields; ++f) {
        cx<T> v[1][SZ::RMAX];
        if constexpr (LDSIN) {
            __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // this wave's loads of field f (and its stores)
            lds_barrier_nofence();                      // and every wave's
            if (j < NR0) {
#pragma unroll
                for (int r = 0; r < R0; ++r) v[0][r] = cx_as<T>(stg[(j + r * NR0) * G + g]);
            }
            lds_barrier_nofence();  // every wave has read the staged tile
            if (f + 1 < a.nfields) issue(f + 1);
        } else if constexpr (PF) {
#pragma unroll
            for (int r = 0; r < SZ::RMAX; ++r) v[0][r] = cx_as<T>(nxt[r]);
            if (f + 1 < a.nfields) load_field(f + 1, nxt);
        } else {
            cx<TI> cur[SZ::RMAX];
            load_field(f, cur);
#pragma unroll
            for (int r = 0; r < SZ::RMAX; ++r) v[0][r] = cx_as<T>(cur[r]);
        }
        // LDSIN: the barrier after the staged tile is read also orders the previous field's last
        // reads of the stage buffer before this field's first write
        transform<T, SZ, 1, true, LP, false, LDSIN, LDSIN, !LDSIN>(par ? bufB : bufA, par ? bufA : bufB, v, j, g, 0, Wl);
        if (ALT) par ^= 1;
        if (col_ok && j < SZ::NRL) {
            cx<T>* o = a.out + (size_t)f * a.ncell + tile0;
            const uint32_t out_lane = (g + jl * rstride) * (uint32_t)sizeof(cx<T>);
            const uint32_t w_lane = q * jl * (uint32_t)sizeof(cx<T>);
#pragma unroll
            for (int r = 0; r < SZ::RL; ++r) {
                // row k2 = j + r NRL: its twiddle W^{q k2} and its output row
                const uint32_t k2u = jw + (uint32_t)(r * SZ::NRL);
                const cx<T> w = LDSIN ? wsl[j + r * SZ::NRL] : load_at(a.Wsplit + q * k2u, w_lane);
                store_at(o + (size_t)k2u * rstride, out_lane, cmul(v[0][r], w));
            }
        }
    }
}

}  // namespace qpd
