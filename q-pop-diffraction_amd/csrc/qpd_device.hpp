// Device building blocks for gfx950: complex arithmetic, fast division by a
// run-time constant, FP64-reduced FP32 sincos, and the LDS-resident
// mixed-radix Stockham FFT used by every pass kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace qpd {

template <typename T>
struct cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// -i * a
template <typename T>
__device__ __forceinline__ cx<T> negi(cx<T> a) { return {a.y, -a.x}; }
// a complex value converted to element type T (exact when T is at least as wide)
template <typename T, typename S>
__device__ __forceinline__ cx<T> cx_as(cx<S> a) { return {(T)a.x, (T)a.y}; }

// Entry i (wave-uniform) of a table that no kernel writes while it runs, read through the
// constant address space, i.e. as a scalar load.  A table entry read as a vector load is
// waited for with the vector-memory counter, which also waits for every vector load issued
// before it (a prefetched input line included); a scalar load waits on its own counter.
template <typename V>
__device__ __forceinline__ V uload(const V* p, int i) {
    return ((const __attribute__((address_space(4))) V*)p)[i];
}

// Element type of the arrays an FP64 kernel reads and writes: with IO32 it computes in FP64
// and keeps its complex fields, P, the phase fractions and the region weight in FP32 (an FP32
// plan whose transforms run in FP64).
template <typename T, bool IO32>
using io_t = std::conditional_t<IO32, float, T>;

// Division by a run-time constant d: q = floor(n / d) for n * d < 2^32.
// m = floor((2^32 - 1) / d) + 1, q = umulhi(n, m); d == 1 is flagged by m == 0.
struct FastDiv {
    uint32_t d;
    uint32_t m;
};

__host__ inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    f.m = (d <= 1) ? 0u : (uint32_t)(0xFFFFFFFFu / d + 1u);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv f) {
    return f.m == 0u ? n : __umulhi(n, f.m);
}

// sin and cos of an FP64 phase t (radians), FP32 results.  The argument is
// reduced to quarter turns in FP64 (r = t * 2/pi, q = rint(r)), so phases of
// thousands of radians keep FP32 accuracy; the remainder |x| <= pi/4 goes
// through the Taylor series of sin (to x^9) and cos (to x^10), whose
// truncation errors (< 2e-9 and < 2e-10) are below FP32 rounding.
__device__ __forceinline__ void sincos_fp64_reduced(double t, float& s, float& c) {
    const double two_over_pi = 0.636619772367581343075535053490057448;
    const double half_pi = 1.570796326794896619231321691639751442;
    double r = t * two_over_pi;
    double q = rint(r);
    float x = (float)((r - q) * half_pi);
    float x2 = x * x;
    float ps = fmaf(x2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = fmaf(x2, ps, 1.0f / 120.0f);
    ps = fmaf(x2, ps, -1.0f / 6.0f);
    ps = fmaf(x2, ps, 1.0f);
    float sx = x * ps;
    float pc = fmaf(x2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = fmaf(x2, pc, -1.0f / 720.0f);
    pc = fmaf(x2, pc, 1.0f / 24.0f);
    pc = fmaf(x2, pc, -0.5f);
    float cx_ = fmaf(x2, pc, 1.0f);
    int quadrant = ((int)q) & 3;
    float s0 = (quadrant & 1) ? cx_ : sx;
    float c0 = (quadrant & 1) ? sx : cx_;
    s = (quadrant & 2) ? -s0 : s0;
    c = ((quadrant + 1) & 2) ? -c0 : c0;
}

// sin and cos of a phase of a few radians given in FP32: quarter-turn reduction in
// FP32 (absolute error about 1e-7 rad for |t| < 4), then the same series.
__device__ __forceinline__ void sincos_small(float t, float& s, float& c) {
    const float r = t * 0.636619772367581343f;
    const float q = rintf(r);
    const float x = fmaf(q, -1.57079637050628662f, t) + q * 4.37113883e-8f;  // t - q pi/2, two-part pi/2
    const float x2 = x * x;
    float ps = fmaf(x2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = fmaf(x2, ps, 1.0f / 120.0f);
    ps = fmaf(x2, ps, -1.0f / 6.0f);
    ps = fmaf(x2, ps, 1.0f);
    const float sx = x * ps;
    float pc = fmaf(x2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = fmaf(x2, pc, -1.0f / 720.0f);
    pc = fmaf(x2, pc, 1.0f / 24.0f);
    pc = fmaf(x2, pc, -0.5f);
    const float cx_ = fmaf(x2, pc, 1.0f);
    const int quadrant = ((int)q) & 3;
    const float s0 = (quadrant & 1) ? cx_ : sx;
    const float c0 = (quadrant & 1) ? sx : cx_;
    s = (quadrant & 2) ? -s0 : s0;
    c = ((quadrant + 1) & 2) ? -c0 : c0;
}

__device__ __forceinline__ void phase_sincos(double t, float& s, float& c) { sincos_fp64_reduced(t, s, c); }
__device__ __forceinline__ void phase_sincos(double t, double& s, double& c) { sincos(t, &s, &c); }

// ---------------------------------------------------------------------------
// 1-D transform plan of one axis: N = prod R[s]; stage s has sub-length Ns
// (product of the earlier radices) and twiddle stride tws = N / (Ns R) into the
// table W[m] = exp(-2 pi i m / N).  The stage list lives in device memory and
// is read with uniform (scalar) loads.
constexpr int kMaxStages = 16;
// Complex values a thread holds during one stage: 16, or 8 in the kRadixPow2 instantiations of
// fft_tile (FP64 only: the same 64 VGPRs as 16 FP32 values; they spilled to scratch with 16).
constexpr int kMaxRegs = 16;
constexpr int kMaxRegsPow2 = 8;
// values a thread holds in the stages of the kernels with sixteen elements a thread (lines
// of more than 4096 points): twice kMaxRegs, so that every radix's butterflies of a line of
// 8192 points fit 512 threads
constexpr int kMaxRegsLong = 32;

struct Stage {
    int R;
    int Ns;
    int tws;
    uint32_t ns_m;  // FastDiv multiplier of Ns
    int NR;         // N / R, read per stage so that nothing stage-specific is hoisted out of the stage loop
    int pad[3];
};

// Host-side description of an axis (the device sees N, nst and the Stage list).
struct AxisFFT {
    int N;
    int nst;
    Stage st[kMaxStages];
};

template <typename T>
__device__ __forceinline__ void dft2(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <typename T>
__device__ __forceinline__ void dft4(cx<T>& v0, cx<T>& v1, cx<T>& v2, cx<T>& v3) {
    cx<T> t0 = v0 + v2, t1 = v0 - v2, t2 = v1 + v3, t3 = negi(v1 - v3);
    v0 = t0 + t2;
    v2 = t0 - t2;
    v1 = t1 + t3;
    v3 = t1 - t3;
}

// 3-point DFT in place (w = exp(-2 pi i / 3) = -1/2 - i sqrt(3)/2).
template <typename T>
__device__ __forceinline__ void dft3(cx<T>& a, cx<T>& b, cx<T>& c) {
    const T h = (T)0.866025403784438646763723170752936183;  // sqrt(3) / 2
    const cx<T> t1 = b + c, t2 = b - c;
    const cx<T> m = {a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y};
    a = a + t1;
    b = {m.x + h * t2.y, m.y - h * t2.x};
    c = {m.x - h * t2.y, m.y + h * t2.x};
}

// 6-point DFT: 3-point DFTs of the even and odd inputs, the odd ones times
// exp(-2 pi i k / 6) (k = 1, 2), joined by a radix-2 step.
template <typename T>
__device__ __forceinline__ void dft6(cx<T>* v) {
    const T h = (T)0.866025403784438646763723170752936183;  // sqrt(3) / 2
    cx<T> e0 = v[0], e1 = v[2], e2 = v[4];
    cx<T> o0 = v[1], o1 = v[3], o2 = v[5];
    dft3(e0, e1, e2);
    dft3(o0, o1, o2);
    o1 = {(T)0.5 * o1.x + h * o1.y, (T)0.5 * o1.y - h * o1.x};
    o2 = {-(T)0.5 * o2.x + h * o2.y, -(T)0.5 * o2.y - h * o2.x};
    v[0] = e0 + o0;
    v[3] = e0 - o0;
    v[1] = e1 + o1;
    v[4] = e1 - o1;
    v[2] = e2 + o2;
    v[5] = e2 - o2;
}

template <typename T>
__device__ __forceinline__ void dft8(cx<T>* v) {
    const T h = (T)0.707106781186547524400844362104849039;  // sqrt(1/2)
    cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    // o_k *= exp(-2 pi i k / 8)
    o1 = {h * (o1.x + o1.y), h * (o1.y - o1.x)};
    o2 = negi(o2);
    o3 = {h * (o3.y - o3.x), -h * (o3.x + o3.y)};
    v[0] = e0 + o0;
    v[4] = e0 - o0;
    v[1] = e1 + o1;
    v[5] = e1 - o1;
    v[2] = e2 + o2;
    v[6] = e2 - o2;
    v[3] = e3 + o3;
    v[7] = e3 - o3;
}

// cos and sin of 2 pi m / R (m = 0..R/2) at compile time, for the odd-radix DFTs in FP32:
// Taylor series in double on |x| <= pi / 2 (x = pi - 2 pi m / R past a quarter turn), then
// rounded to float.  The host's twiddle table rounds cosl / sinl to double and then to
// float, so the two agree unless a value falls within a double ulp of a float rounding
// boundary.  As literals they cost no registers and no LDS reads.
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}
constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}

template <int R>
struct OddRoots {
    float c[R / 2 + 1], s[R / 2 + 1];
    constexpr OddRoots() : c(), s() {
        constexpr double pi = 3.14159265358979323846264338327950288;
        for (int m = 0; m <= R / 2; ++m) {
            const bool past = 4 * m > R;  // angle above pi / 2
            const double x = past ? pi * (double)(R - 2 * m) / (double)R : 2.0 * pi * (double)m / (double)R;
            c[m] = (float)(past ? -taylor_cos(x) : taylor_cos(x));
            s[m] = (float)taylor_sin(x);
        }
    }
};

template <int R>
__device__ constexpr OddRoots<R> kOddRoots{};

// The same constants for the FP64 DFTs, evaluated in double-double arithmetic (a value is an
// unevaluated sum hi + lo of two doubles, about 106 bits; Dekker's product split, Knuth's
// two-sum) and rounded to double: the double nearest cos, sin(2 pi m / R), which is the value
// of the host's twiddle table (cosl / sinl rounded to double).  As literals they take no
// vector registers: an FP64 radix-13 stage otherwise holds its twelve in 24 VGPRs.
struct DD {
    double hi, lo;
};
constexpr DD dd_two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
constexpr DD dd_fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
constexpr DD dd_split(double a) {
    const double t = 134217729.0 * a;  // 2^27 + 1
    const double hi = t - (t - a);
    return {hi, a - hi};
}
constexpr DD dd_two_prod(double a, double b) {
    const double p = a * b;
    const DD as = dd_split(a), bs = dd_split(b);
    return {p, ((as.hi * bs.hi - p) + as.hi * bs.lo + as.lo * bs.hi) + as.lo * bs.lo};
}
constexpr DD dd_add(DD a, DD b) {
    DD s = dd_two_sum(a.hi, b.hi);
    const DD t = dd_two_sum(a.lo, b.lo);
    s = dd_fast_two_sum(s.hi, s.lo + t.hi);
    return dd_fast_two_sum(s.hi, s.lo + t.lo);
}
constexpr DD dd_mul(DD a, DD b) {
    const DD p = dd_two_prod(a.hi, b.hi);
    return dd_fast_two_sum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}
constexpr DD dd_div(DD a, double b) {  // b a small integer
    const double q1 = a.hi / b;
    const DD p = dd_two_prod(q1, b);
    const DD r = dd_add(a, {-p.hi, -p.lo});
    return dd_fast_two_sum(q1, r.hi / b);
}
// sin (odd) or cos of x, |x| <= pi / 2, by its Taylor series to 27 terms
constexpr double dd_taylor(DD x, bool odd) {
    const DD x2 = dd_mul(x, x);
    DD term = odd ? x : DD{1.0, 0.0};
    DD sum = term;
    for (int k = 1; k < 28; ++k) {
        const double d = odd ? (2.0 * k) * (2.0 * k + 1.0) : (2.0 * k - 1.0) * (2.0 * k);
        term = dd_div(dd_mul(term, {-x2.hi, -x2.lo}), d);
        sum = dd_add(sum, term);
    }
    return sum.hi + sum.lo;
}

template <int R>
struct OddRoots64 {
    double c[R / 2 + 1], s[R / 2 + 1];
    constexpr OddRoots64() : c(), s() {
        constexpr DD pi = {3.141592653589793116, 1.2246467991473532e-16};
        for (int m = 0; m <= R / 2; ++m) {
            const bool past = 4 * m > R;  // angle above pi / 2
            const DD x = dd_div(dd_mul(pi, {(double)(past ? R - 2 * m : 2 * m), 0.0}), (double)R);
            c[m] = past ? -dd_taylor(x, false) : dd_taylor(x, false);
            s[m] = dd_taylor(x, true);
        }
    }
};

template <int R>
__device__ constexpr OddRoots64<R> kOddRoots64{};

// Odd-length DFT y[t] = sum_s v[s] exp(-2 pi i t s / R) using the symmetric pairs
// a_s = v[s] + v[R-s], b_s = v[s] - v[R-s], with cos, sin(2 pi m / R) the compile-time
// constants kOddRoots (FP32) or kOddRoots64 (FP64).
template <int R, typename T>
__device__ __forceinline__ void dft_odd(cx<T>* v) {
    constexpr int H = (R - 1) / 2;
    cx<T> a[H + 1], b[H + 1];
    cx<T> y0 = v[0];
#pragma unroll
    for (int s = 1; s <= H; ++s) {
        a[s] = v[s] + v[R - s];
        b[s] = v[s] - v[R - s];
        y0 = y0 + a[s];
    }
    cx<T> out[R];
    out[0] = y0;
#pragma unroll
    for (int t = 1; t <= H; ++t) {
        cx<T> A = v[0];
        cx<T> B = {(T)0, (T)0};
#pragma unroll
        for (int s = 1; s <= H; ++s) {
            const int m = (t * s) % R;
            const int mm = m <= H ? m : R - m;
            constexpr bool f32 = std::is_same<T, float>::value;
            const T c = f32 ? (T)kOddRoots<R>.c[mm] : (T)kOddRoots64<R>.c[mm];
            const T s1 = f32 ? (T)kOddRoots<R>.s[mm] : (T)kOddRoots64<R>.s[mm];
            const T sn = m <= H ? s1 : -s1;
            A.x += a[s].x * c;
            A.y += a[s].y * c;
            B.x += b[s].x * sn;
            B.y += b[s].y * sn;
        }
        out[t] = {A.x + B.y, A.y - B.x};
        out[R - t] = {A.x - B.y, A.y + B.x};
    }
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = out[t];
}

template <int R, typename T>
__device__ __forceinline__ void dft(cx<T>* v) {
    if constexpr (R == 2) {
        dft2(v);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 6) {
        dft6(v);
    } else if constexpr (R == 8) {
        dft8(v);
    } else {
        dft_odd<R>(v);
    }
}

// One Stockham stage, in place in LDS.  The tile holds L lines of N points,
// point j of line l at tile[j * stride + l]; lines [0, Lc) are transformed.
// Every thread first loads all its butterflies (at most MR values), the
// block synchronises, then the butterflies are twiddled, transformed and
// written back in autosort order.
template <int R, typename T, int MR = kMaxRegs>
__device__ __forceinline__ void fft_stage(cx<T>* __restrict__ tile, const cx<T>* __restrict__ W, int NR,
                                          int stride, int Lc, FastDiv divLc, int Ns, FastDiv divNs, int tws,
                                          int tid, int nthr) {
    constexpr int MB = (MR / R) > 0 ? (MR / R) : 1;
    const int total = NR * Lc;
    cx<T> v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int e = tid + b * nthr;
        if (e < total) {
            const int j = (int)fdiv((uint32_t)e, divLc);
            const int line = e - j * Lc;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = tile[(j + r * NR) * stride + line];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int e = tid + b * nthr;
        if (e < total) {
            const int j = (int)fdiv((uint32_t)e, divLc);
            const int line = e - j * Lc;
            const int q = (int)fdiv((uint32_t)j, divNs);
            const int k = j - q * Ns;
            if (Ns > 1) {
                const int base = k * tws;
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], W[r * base]);
            }
            dft<R>(v[b]);
            const int o = q * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) tile[(o + r * Ns) * stride + line] = v[b][r];
        }
    }
    __syncthreads();
}

// A stage of prime radix R above 17 (the leftover factors of any length): each output is
// formed directly, y_t = sum_r x[j + r NR] w^(r k tws) w^(r t NR) with w = exp(-2 pi i / N)
// (the stage twiddle and the DFT_R root in one table entry, index r (k tws + t NR) mod N),
// O(R) work an output.  Output pos = q Ns R + t Ns + k of butterfly j = q Ns + k, as in
// fft_stage.  The stage's inputs are first copied to the second half of the LDS (the host
// gives kernels of such axes twice the tile, N * stride points after it), so that each
// thread forms and stores its outputs one at a time, with no outputs held in registers
// across the barrier (which would cost every kernel that compiles this stage registers).
// The sum runs in blocks of 16 terms (partial sums added last), which keeps the FP32
// rounding of a long sum near that of a tree.
template <typename T>
__device__ __forceinline__ void fft_stage_direct(cx<T>* __restrict__ tile, const cx<T>* __restrict__ W, int R, int NR,
                                                  int stride, int Lc, FastDiv divLc, int Ns, FastDiv divNs, int tws,
                                                  int tid, int nthr) {
    const int N = NR * R;
    const int total = N * Lc;
    cx<T>* src = tile + (size_t)N * stride;
    for (int e = tid; e < total; e += nthr) {
        const int pos = (int)fdiv((uint32_t)e, divLc);
        const int at = pos * stride + (e - pos * Lc);
        src[at] = tile[at];
    }
    __syncthreads();
    for (int e = tid; e < total; e += nthr) {
        const int pos = (int)fdiv((uint32_t)e, divLc);
        const int line = e - pos * Lc;
        const int qt = (int)fdiv((uint32_t)pos, divNs);  // q R + t
        const int k = pos - qt * Ns;
        const int q = qt / R, t = qt - q * R;
        const cx<T>* x = src + (q * Ns + k) * stride + line;
        const int step = k * tws + t * NR;  // < N
        const int xstep = NR * stride;
        int idx = 0;
        cx<T> acc = {(T)0, (T)0};
        for (int r0 = 0; r0 < R; r0 += 16) {
            const int r1 = min(R, r0 + 16);
            cx<T> part = {(T)0, (T)0};
            for (int r = r0; r < r1; ++r) {
                part = part + cmul(x[r * xstep], W[idx]);
                idx += step;
                if (idx >= N) idx -= N;
            }
            acc = acc + part;
        }
        tile[pos * stride + line] = acc;
    }
    __syncthreads();
}

// Radix sets of the generic kernels: every radix (with the direct stages of primes above 17),
// radices 2, 4 and 8 only, or the radices up to 8 (2, 3, 4, 5, 7, 8).  A kernel's registers
// are those of its largest stage (the FP64 radix-17 stage alone holds about 200 VGPRs), so
// FP64 axes of small radices run instantiations without the large ones.
constexpr int kRadixAll = 0, kRadixPow2 = 1, kRadixSmall = 2;

// Forward transform of lines [0, Lc) of an LDS tile (see fft_stage), natural order in and out,
// for an axis whose radices are all in the set RS (the host picks RS per axis).
template <typename T, int RS = kRadixAll, int MR = kMaxRegs>
__device__ __forceinline__ void fft_tile(cx<T>* tile, int nst, const Stage* __restrict__ stages,
                                         const cx<T>* __restrict__ W, int stride, int Lc, FastDiv divLc, int tid,
                                         int nthr) {
    for (int s = 0; s < nst; ++s) {
        const Stage sd = stages[s];
        const FastDiv dn = {(uint32_t)sd.Ns, sd.ns_m};
        const int Ns = sd.Ns, tws = sd.tws, NR = sd.NR;
        if constexpr (RS == kRadixPow2) {
            switch (sd.R) {
                case 2: fft_stage<2, T, kMaxRegsPow2>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                case 4: fft_stage<4, T, kMaxRegsPow2>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                case 8: fft_stage<8, T, kMaxRegsPow2>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                default: break;  // the host picks this set only for axes of these radices
            }
        } else {
            switch (sd.R) {
                case 2: fft_stage<2, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                case 3: fft_stage<3, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                case 4: fft_stage<4, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                case 5: fft_stage<5, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                case 7: fft_stage<7, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                case 8: fft_stage<8, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                default:
                    if constexpr (RS == kRadixAll) {
                        switch (sd.R) {
                            case 11: fft_stage<11, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                            case 13: fft_stage<13, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                            case 17: fft_stage<17, T, MR>(tile, W, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr); break;
                            default:  // primes above 17
                                fft_stage_direct<T>(tile, W, sd.R, NR, stride, Lc, divLc, Ns, dn, tws, tid, nthr);
                                break;
                        }
                    }
                    break;  // kRadixSmall: the host picks this set only for axes of radices up to 8
            }
        }
    }
}

}  // namespace qpd
