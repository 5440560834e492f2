// libqpd: C ABI of the MI355X kinematic-diffraction path (see include/qpd.h).
//
// Host side: plan creation (factorisation of each axis, FP64 twiddle / wavenumber /
// coefficient tables rounded once to the working precision, atom classes, launch
// shapes) and the entry points that enqueue the three pass kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "qpd.h"
#include "qpd_kernels.hpp"
#include "qpd_fast.hpp"

using namespace qpd;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define QPD_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t err_ = (call);                                                                   \
        if (err_ != hipSuccess)                                                                     \
            return fail(QPD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(err_));           \
    } while (0)

// ---------------------------------------------------------------------------
// Axis factorisation: radix 8 while possible, then 4 or 2, then the odd primes.
bool factorise(int64_t n, std::vector<int>& radices) {
    radices.clear();
    if (n < 1) return false;
    int64_t m = n;
    while (m % 8 == 0) {
        radices.push_back(8);
        m /= 8;
    }
    if (m % 4 == 0) {
        radices.push_back(4);
        m /= 4;
    }
    if (m % 2 == 0) {
        radices.push_back(2);
        m /= 2;
    }
    for (int p : {3, 5, 7, 11, 13, 17}) {
        while (m % p == 0) {
            radices.push_back(p);
            m /= p;
        }
    }
    return m == 1;
}

bool make_axis(int64_t N, AxisFFT& ax) {
    std::vector<int> r;
    if (!factorise(N, r) || (int)r.size() > kMaxStages || N > 65536) return false;
    std::memset(&ax, 0, sizeof(ax));
    ax.N = (int)N;
    ax.nst = (int)r.size();
    int ns = 1;
    for (int s = 0; s < ax.nst; ++s) {
        ax.st[s].R = r[s];
        ax.st[s].Ns = ns;
        ax.st[s].tws = (int)(N / ((int64_t)ns * r[s]));
        ax.st[s].ns_m = make_fastdiv((uint32_t)ns).m;
        ax.st[s].NR = (int)(N / r[s]);
        ns *= r[s];
    }
    return true;
}

// exp(-i theta) with theta = 2 pi num / den, evaluated in extended precision.
void unit_root(int64_t num, int64_t den, double& re, double& im) {
    const long double two_pi = 6.283185307179586476925286766559005768L;
    long double t = two_pi * (long double)(num % den) / (long double)den;
    re = (double)cosl(t);
    im = (double)-sinl(t);
}

// 2 pi fftfreq(N, h)[i], nm^-1 (kernels.cuh:15-18; numpy.fft.fftfreq).
double wavenumber(int64_t i, int64_t N, double h) {
    int64_t m = i <= (N - 1) / 2 ? i : i - N;
    return 2.0 * M_PI * (double)m / ((double)N * h);
}

// numpy.sinc(k h / 2 pi) = sin(k h / 2) / (k h / 2) (diffraction.py:85-87).
double shape_factor(double k, double h) {
    long double x = (long double)k * h * 0.5L;
    if (x == 0.0L) return 1.0;
    return (double)(sinl(x) / x);
}

// The compile-time-length kernels are used unless QPD_GENERIC_FFT=1 (for comparison runs).
bool fast_enabled() {
    static int v = -1;
    if (v < 0) {
        const char* e = std::getenv("QPD_GENERIC_FFT");
        v = (e && e[0] == '1') ? 0 : 1;
    }
    return v == 1;
}

int max_lds_bytes() {
    static int cached = -1;
    if (cached < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            cached = v;
        else
            cached = 65536;
    }
    return cached;
}

// ---------------------------------------------------------------------------
// Launch shape of one pass.
struct Launch {
    int fast;        // 1: compile-time-length kernel of qpd_fast.hpp (FP32), 0: generic
    int contiguous;  // 1: pass_contig, 0: pass_strided
    int nl;          // rows (contig) or TX (strided) per block
    int nf;          // inputs per LDS fill
    int nthr;
    int ept;         // elements per thread, 4 / 8 / 16
    size_t lds;
    int64_t blocks;
};

bool stage_fits(const AxisFFT& ax, int lines, int nthr) {
    for (int s = 0; s < ax.nst; ++s) {
        const int R = ax.st[s].R;
        const int mb = std::max(1, kMaxRegs / R);
        const int64_t total = (int64_t)(ax.N / R) * lines;
        if ((total + nthr - 1) / nthr > mb) return false;
    }
    return true;
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Pick (lines per field, inputs per fill, threads, EPT) for a pass over lines of length N.
bool choose_launch(const AxisFFT& ax, int nin, bool contiguous, size_t csize, int64_t n_lines_total,
                   int lds_soft, int lds_hard, int ept_max, Launch& L) {
    const int N = ax.N;
    std::vector<int> nfs;
    for (int nf : {4, 3, 2, 1})
        if (nf <= nin) nfs.push_back(nf);
    std::vector<int> per;  // lines per field candidates
    if (contiguous) {
        for (int nl = 64; nl >= 1; nl /= 2) per.push_back(nl);
    } else {
        const int tx_pref = csize == 8 ? 16 : 8;
        for (int tx = tx_pref; tx >= 1; tx /= 2) per.push_back(tx);
    }
    // contiguous passes prefer holding all inputs of a chunk at once; strided passes
    // prefer wide column tiles (coalesced rows) over more inputs per fill
    std::vector<std::pair<int, int>> order;
    if (contiguous) {
        for (int nf : nfs)
            for (int p : per) order.push_back({nf, p});
    } else {
        for (int p : per)
            for (int nf : nfs) order.push_back({nf, p});
    }
    for (int pass = 0; pass < 2; ++pass) {
        const int cap = pass == 0 ? lds_soft : lds_hard;
        for (auto [nf, p] : order) {
            {
                if (contiguous && p > 1 && (int64_t)p > n_lines_total) continue;
                const int64_t elems = (int64_t)p * N;      // one field's tile
                if (contiguous && elems > 4096 && p > 1) continue;
                const size_t lds = (size_t)nf * elems * csize;
                if (lds > (size_t)cap) continue;
                // threads: aim at ~ept_max/2 elements per thread, at most 512
                int nthr = round_up((int)std::min<int64_t>(512, (elems + ept_max / 2 - 1) / (ept_max / 2)), 64);
                nthr = std::max(64, std::min(512, nthr));
                while (nthr < 512 && (elems + nthr - 1) / nthr > ept_max) nthr += 64;
                if ((elems + nthr - 1) / nthr > ept_max) continue;
                while (nthr < 512 && !stage_fits(ax, nf * p, nthr)) nthr += 64;
                if (!stage_fits(ax, nf * p, nthr)) continue;
                L.contiguous = contiguous;
                L.nl = p;
                L.nf = nf;
                L.nthr = nthr;
                const int need = (int)((elems + nthr - 1) / nthr);
                L.ept = need <= 4 ? 4 : 8;
                if (L.ept > ept_max) continue;
                L.lds = lds;
                return true;
            }
        }
    }
    return false;
}

}  // namespace

// ---------------------------------------------------------------------------
struct qpd_plan {
    int precision;
    int64_t dims[3];
    int axis[3];
    int64_t ncell;
    int n_phase;
    int n_atom;
    int nY, nZ;  // classes after pass 0 and pass 1
    size_t csize;
    AxisFFT fft[3];
    Launch launch[3];
    int ngroups[3];
    // device tables
    void* dev = nullptr;          // one allocation holding every table below
    const void* W[3];
    const Stage* stages[3];
    const void* k[3];
    const void* coef[3];
    const int* chunk_src[3];
    const int* chunk_coef[3];
    const int* group_first[3];
    const void* atomtab;
    double qC[3];
};

namespace {

struct TableBuilder {
    std::vector<unsigned char> bytes;
    size_t add(const void* p, size_t n) {
        size_t off = (bytes.size() + 255) / 256 * 256;
        bytes.resize(off + n);
        if (n) std::memcpy(bytes.data() + off, p, n);
        return off;
    }
    template <typename T>
    size_t add_real(const std::vector<double>& v) {
        std::vector<T> t(v.begin(), v.end());
        return add(t.data(), t.size() * sizeof(T));
    }
    template <typename T>
    size_t add_cplx(const std::vector<double>& re, const std::vector<double>& im) {
        std::vector<T> t(2 * re.size());
        for (size_t i = 0; i < re.size(); ++i) {
            t[2 * i] = (T)re[i];
            t[2 * i + 1] = (T)im[i];
        }
        return add(t.data(), t.size() * sizeof(T));
    }
};

template <typename T>
int build_tables(qpd_plan* p, const qpd_grid* g, const qpd_crystal* c) {
    TableBuilder tb;
    size_t offW[3], offK[3], offC[3], offSrc[3], offCoef[3], offGrp[3], offSt[3], offAtom;
    const int64_t n = p->ncell;
    const double h_of[3] = {g->h[0], g->h[1], g->h[2]};

    // atom classes: pass-0 groups by equal (R[ax1], R[ax2]); pass-1 groups by equal R[ax2]
    const int ax0 = g->axis[0], ax1 = g->axis[1], ax2 = g->axis[2];
    std::vector<int> ycls(c->n_atom), zcls;
    std::vector<std::pair<double, double>> ykeys;
    std::vector<double> zkeys;
    for (int a = 0; a < c->n_atom; ++a) {
        std::pair<double, double> key(c->R[3 * a + ax1], c->R[3 * a + ax2]);
        auto it = std::find(ykeys.begin(), ykeys.end(), key);
        if (it == ykeys.end()) {
            ykeys.push_back(key);
            ycls[a] = (int)ykeys.size() - 1;
        } else {
            ycls[a] = (int)(it - ykeys.begin());
        }
    }
    zcls.resize(ykeys.size());
    for (size_t y = 0; y < ykeys.size(); ++y) {
        auto it = std::find(zkeys.begin(), zkeys.end(), ykeys[y].second);
        if (it == zkeys.end()) {
            zkeys.push_back(ykeys[y].second);
            zcls[y] = (int)zkeys.size() - 1;
        } else {
            zcls[y] = (int)(it - zkeys.begin());
        }
    }
    p->nY = (int)ykeys.size();
    p->nZ = (int)zkeys.size();

    for (int d = 0; d < 3; ++d) {
        const int64_t N = g->dims[d];
        const double h = h_of[g->axis[d]];
        std::vector<double> wr(N), wi(N), kk(N);
        for (int64_t m = 0; m < N; ++m) {
            unit_root(m, N, wr[m], wi[m]);
            kk[m] = wavenumber(m, N, h);
        }
        offW[d] = tb.add_cplx<T>(wr, wi);
        offK[d] = tb.add_real<T>(kk);
        offSt[d] = tb.add(p->fft[d].st, sizeof(Stage) * kMaxStages);

        // coefficient tables and chunk / group lists
        std::vector<double> cr, ci;
        std::vector<int> src, coef, first;
        auto add_table = [&](double R_d, std::complex<double> scale) {
            for (int64_t m = 0; m < N; ++m) {
                const double kv = kk[m];
                const double s = shape_factor(kv, h);
                const std::complex<double> ph = std::polar(1.0, -kv * R_d);
                const std::complex<double> v = scale * ph * s;
                cr.push_back(v.real());
                ci.push_back(v.imag());
            }
        };
        if (d == 0) {
            // pass 0: group = y-class, chunk = atom; c_a = exp(-i G.R_a) / (V n) exp(-i k R_a,ax0) s(k)
            for (int y = 0; y < p->nY; ++y) {
                first.push_back((int)src.size());
                for (int a = 0; a < c->n_atom; ++a) {
                    if (ycls[a] != y) continue;
                    src.push_back(a);
                    coef.push_back(a);
                }
            }
            first.push_back((int)src.size());
            for (int a = 0; a < c->n_atom; ++a) {
                const double gr = c->G[0] * c->R[3 * a] + c->G[1] * c->R[3 * a + 1] + c->G[2] * c->R[3 * a + 2];
                add_table(c->R[3 * a + ax0], std::polar(1.0, -gr) / (c->V * (double)n));
            }
        } else if (d == 1) {
            // pass 1: group = z-class, chunk = y-class; c = exp(-i k R_ax1) s(k)
            for (int z = 0; z < p->nZ; ++z) {
                first.push_back((int)src.size());
                for (int y = 0; y < p->nY; ++y)
                    if (zcls[y] == z) {
                        src.push_back(y);
                        coef.push_back(y);
                    }
            }
            first.push_back((int)src.size());
            for (int y = 0; y < p->nY; ++y) add_table(ykeys[y].first, 1.0);
        } else {
            // pass 2: one group, chunk = z-class; c = exp(-i k R_ax2) s(k)
            first.push_back(0);
            for (int z = 0; z < p->nZ; ++z) {
                src.push_back(z);
                coef.push_back(z);
            }
            first.push_back((int)src.size());
            for (int z = 0; z < p->nZ; ++z) add_table(zkeys[z], 1.0);
        }
        offC[d] = tb.add_cplx<T>(cr, ci);
        offSrc[d] = tb.add(src.data(), src.size() * sizeof(int));
        offCoef[d] = tb.add(coef.data(), coef.size() * sizeof(int));
        offGrp[d] = tb.add(first.data(), first.size() * sizeof(int));
        p->ngroups[d] = (int)first.size() - 1;
    }

    // atom table [a][p]: Re f, Im f, su[d][k] (d physical direction, k order-parameter component)
    std::vector<double> at;
    for (int a = 0; a < c->n_atom; ++a)
        for (int ph = 0; ph < c->n_phase; ++ph) {
            at.push_back(c->f_re[ph * c->n_atom + a]);
            at.push_back(c->f_im[ph * c->n_atom + a]);
            for (int d = 0; d < 3; ++d)
                for (int kk = 0; kk < 3; ++kk)
                    at.push_back(c->su[((d * c->n_struc + kk) * c->n_phase + ph) * c->n_atom + a]);
            for (int kk = 0; kk < 3; ++kk) {  // g_k = sum_d qC_d su[d][k], in FP64
                double gk = 0.0;
                for (int d = 0; d < 3; ++d) gk += c->qC[d] * c->su[((d * c->n_struc + kk) * c->n_phase + ph) * c->n_atom + a];
                at.push_back(gk);
            }
        }
    offAtom = tb.add_real<T>(at);

    QPD_HIP(hipMalloc(&p->dev, tb.bytes.size()));
    QPD_HIP(hipMemcpy(p->dev, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice));
    unsigned char* base = static_cast<unsigned char*>(p->dev);
    for (int d = 0; d < 3; ++d) {
        p->W[d] = base + offW[d];
        p->stages[d] = reinterpret_cast<const Stage*>(base + offSt[d]);
        p->k[d] = base + offK[d];
        p->coef[d] = base + offC[d];
        p->chunk_src[d] = reinterpret_cast<const int*>(base + offSrc[d]);
        p->chunk_coef[d] = reinterpret_cast<const int*>(base + offCoef[d]);
        p->group_first[d] = reinterpret_cast<const int*>(base + offGrp[d]);
    }
    p->atomtab = base + offAtom;
    return QPD_OK;
}

template <typename T>
void fill_common(PassArgs<T>& a, const AxisFFT& ax, const Stage* stages_dev, const Launch& L, const int64_t dims[3]) {
    std::memset(&a, 0, sizeof(a));
    a.N = ax.N;
    a.nst = ax.nst;
    a.stages = stages_dev;
    a.n0 = (uint32_t)dims[0];
    a.n1 = (uint32_t)dims[1];
    a.n2 = (uint32_t)dims[2];
    a.ncell = (uint32_t)(dims[0] * dims[1] * dims[2]);
    a.nf = L.nf;
    a.nl = L.nl;
    a.divN = make_fastdiv((uint32_t)ax.N);
    a.divL1 = make_fastdiv((uint32_t)L.nl);
    a.divL2 = make_fastdiv((uint32_t)(2 * L.nl));
    a.divL3 = make_fastdiv((uint32_t)(3 * L.nl));
    a.divL4 = make_fastdiv((uint32_t)(4 * L.nl));
    int lg = 0;
    while ((1 << lg) < L.nl) ++lg;
    a.log2TX = lg;
}

// Raise a kernel's dynamic LDS limit once per (kernel, device) and size, so that launches
// recorded into a graph make no attribute calls after the first eager launch.
hipError_t allow_lds(const void* kernel, int lds) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> granted;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    int& have = granted[{kernel, dev}];
    if (have >= lds) return hipSuccess;
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e == hipSuccess) have = lds;
    return e;
}

template <typename A>
hipError_t launch_one(void (*kernel)(A), const Launch& L, hipStream_t st, const A& a) {
    if (L.lds > 65536) {
        hipError_t e = allow_lds((const void*)kernel, (int)L.lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)L.blocks), dim3((unsigned)L.nthr), L.lds, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Compile-time-length kernels (FP32).  Roles: 0 contiguous + source (4 inputs, 3
// outputs), 1 contiguous plain transform, 2 strided 3 -> 2, 3 strided 2 -> intensity,
// 4 strided plain transform.  G = rows or columns per block.
constexpr int kMaxAtomTab = 1024;  // atom-table entries the source pass keeps in LDS
using S512 = Sizes<512, 8, 8, 8>;
using S256 = Sizes<256, 4, 8, 8>;
using S128 = Sizes<128, 4, 4, 8>;
using S104 = Sizes<104, 8, 13>;
using S64 = Sizes<64, 8, 8>;

constexpr int fast_g(int role, int tpl) {
    // source pass: 4 G TPL threads, about 256; plain contiguous: about 256 threads;
    // strided: 16 columns (128-byte rows of FP32 complex) when the block stays within
    // 512 threads, else 8
    return role == 0 ? (tpl >= 64 ? 1 : 64 / tpl)
                     : role == 1 ? (tpl >= 256 ? 1 : 256 / tpl) : (16 * tpl <= 512 ? 16 : 8);
}

template <class SZ>
bool fast_shape_sz(int role, int atomtab_entries, Launch& L) {
    const int G = fast_g(role, SZ::TPL);
    const int F = role == 0 ? 4 : 1;  // inputs held in the tile at once
    L.fast = 1;
    L.contiguous = role <= 1;
    L.nl = G;
    L.nf = F;
    L.nthr = (role == 0 ? 4 : 1) * G * SZ::TPL;
    L.ept = 0;
    // two stage buffers + the twiddle table (+ the atom table of the source pass)
    const size_t tile = role <= 1 ? (size_t)F * G * SZ::PITCH : (size_t)SZ::N * (F * G + 1);
    L.lds = (2 * tile + SZ::N) * 8 + (role == 0 ? 4 * atomtab_entries : 0);
    return L.nthr <= 1024 && L.lds <= 160 * 1024;
}

bool fast_shape(int64_t N, int role, int atomtab_entries, Launch& L) {
    switch (N) {
        case 512: return fast_shape_sz<S512>(role, atomtab_entries, L);
        case 256: return fast_shape_sz<S256>(role, atomtab_entries, L);
        case 128: return fast_shape_sz<S128>(role, atomtab_entries, L);
        case 104: return fast_shape_sz<S104>(role, atomtab_entries, L);
        case 64: return fast_shape_sz<S64>(role, atomtab_entries, L);
        default: return false;
    }
}

template <class SZ>
hipError_t launch_fast_sz(int role, const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    constexpr int T = SZ::TPL;
    switch (role) {
        case 0: return launch_one(fast_source<float, SZ, fast_g(0, T)>, L, st, a);
        case 1: return launch_one(fast_contig<float, SZ, fast_g(1, T), 1, false>, L, st, a);
        case 2: return launch_one(fast_strided<float, SZ, fast_g(2, T), 3, false>, L, st, a);
        case 3: return launch_one(fast_strided<float, SZ, fast_g(3, T), 2, true>, L, st, a);
        case 4: return launch_one(fast_strided<float, SZ, fast_g(4, T), 1, false>, L, st, a);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fast(int N, int role, const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    switch (N) {
        case 512: return launch_fast_sz<S512>(role, L, a, st);
        case 256: return launch_fast_sz<S256>(role, L, a, st);
        case 128: return launch_fast_sz<S128>(role, L, a, st);
        case 104: return launch_fast_sz<S104>(role, L, a, st);
        case 64: return launch_fast_sz<S64>(role, L, a, st);
        default: return hipErrorInvalidValue;
    }
}

// Kernel instance for a pass: contiguous passes compute the source (3 outputs) or
// transform fields (1 output); strided passes have 2 outputs, or 1 when they
// write the intensity or run a plain transform.
template <typename T>
hipError_t launch_pass(const Launch& L, const PassArgs<T>& a, bool source, bool intensity, hipStream_t st) {
    if (L.fast) {
        if constexpr (std::is_same<T, float>::value) {
            const int role = L.contiguous ? (source ? 0 : 1) : (intensity ? 3 : (a.nout == 2 ? 2 : 4));
            return launch_fast(a.N, role, L, a, st);
        }
        return hipErrorInvalidValue;
    }
    if (L.contiguous) {
        if (source) {
            if (a.nout != 3 || L.ept > 4) return hipErrorInvalidValue;
            return launch_one(pass_contig<T, 4, 3, true>, L, st, a);
        }
        if (a.nout != 1 || L.ept > 8) return hipErrorInvalidValue;
        return L.ept <= 4 ? launch_one(pass_contig<T, 4, 1, false>, L, st, a) : launch_one(pass_contig<T, 8, 1, false>, L, st, a);
    }
    if (L.ept > 8) return hipErrorInvalidValue;
    if (intensity) {
        if (a.nout != 1) return hipErrorInvalidValue;
        return launch_one(pass_strided<T, 8, 1, true>, L, st, a);
    }
    if (a.nout == 2) return launch_one(pass_strided<T, 8, 2, false>, L, st, a);
    if (a.nout != 1) return hipErrorInvalidValue;
    return launch_one(pass_strided<T, 8, 1, false>, L, st, a);
}

// Launch shapes of the three passes of a plan with the given inputs per chunk.
int plan_launches(const AxisFFT fft[3], const int64_t dims[3], const int nin[3], size_t csize, Launch out[3],
                  bool intensity_last = false, int atomtab_entries = 0) {
    const int hard = max_lds_bytes();
    const int soft = std::min(hard, 65536);
    for (int d = 0; d < 3; ++d) {
        const bool contig = d == 0;
        const int64_t lines = contig ? dims[1] * dims[2] : 0;
        if (csize == 8 && fast_enabled()) {
            int role = -1;
            if (contig)
                role = (nin[d] == 4 && atomtab_entries <= kMaxAtomTab) ? 0 : (nin[d] == 1 ? 1 : -1);
            else if (nin[d] == 3)
                role = 2;
            else if (nin[d] == 2 && intensity_last && d == 2)
                role = 3;
            else if (nin[d] == 1)
                role = 4;
            if (role >= 0 && fast_shape(dims[d], role, atomtab_entries, out[d]) && (int)out[d].lds <= hard) {
                if (contig) {
                    out[d].blocks = (lines + out[d].nl - 1) / out[d].nl;
                } else {
                    const int64_t tiles0 = (dims[0] + out[d].nl - 1) / out[d].nl;
                    out[d].blocks = tiles0 * (d == 1 ? dims[2] : dims[1]);
                }
                continue;
            }
        }
        out[d].fast = 0;
        const int ept_max = contig ? (nin[d] > 1 ? 4 : 8) : 8;
        if (!choose_launch(fft[d], nin[d], contig, csize, lines, soft, hard, ept_max, out[d]))
            return fail(QPD_ERR_UNSUPPORTED, "no launch shape fits dim " + std::to_string(d) + " of length " +
                                                 std::to_string(dims[d]));
        if (contig) {
            out[d].blocks = (lines + out[d].nl - 1) / out[d].nl;
        } else {
            const int64_t tiles0 = (dims[0] + out[d].nl - 1) / out[d].nl;
            const int64_t outer = d == 1 ? dims[2] : dims[1];
            out[d].blocks = tiles0 * outer;
        }
        if (out[d].blocks > 0x7fffffffLL) return fail(QPD_ERR_UNSUPPORTED, "grid too large");
    }
    return QPD_OK;
}

template <typename T>
void set_strided(PassArgs<T>& a, int d, const int64_t dims[3]) {
    a.dim = d;
    a.ntiles0 = (uint32_t)((dims[0] + (1 << a.log2TX) - 1) >> a.log2TX);
    if (d == 1) {
        a.S = (uint32_t)dims[0];
        a.OS = (uint32_t)(dims[0] * dims[1]);
    } else {
        a.S = (uint32_t)(dims[0] * dims[1]);
        a.OS = (uint32_t)dims[0];
    }
}

template <typename T>
int run_intensity(qpd_plan* p, const qpd_fields* f, double* I, void* work, hipStream_t st, int only = -1) {
    const int64_t n = p->ncell;
    cx<T>* w1 = static_cast<cx<T>*>(work);
    cx<T>* w2 = w1 + (size_t)3 * p->nY * n;
    for (int d = 0; d < 3; ++d) {
        if (only >= 0 && d != only) continue;
        PassArgs<T> a;
        fill_common(a, p->fft[d], p->stages[d], p->launch[d], p->dims);
        a.W = static_cast<const cx<T>*>(p->W[d]);
        a.coef = static_cast<const cx<T>*>(p->coef[d]);
        a.k = static_cast<const T*>(p->k[d]);
        a.chunk_src = p->chunk_src[d];
        a.chunk_coef = p->chunk_coef[d];
        a.group_first = p->group_first[d];
        a.ngroups = p->ngroups[d];
        a.nin = 4 - d;
        a.nout = 3 - d;
        if (d == 0) {
            a.out = w1;
            a.P = static_cast<const T*>(f->P);
            a.U = f->U;
            a.phase = f->phase;
            a.fraction = static_cast<const T*>(f->fraction);
            a.region = static_cast<const T*>(f->region);
            a.atomtab = static_cast<const T*>(p->atomtab);
            a.qC0 = p->qC[0];
            a.qC1 = p->qC[1];
            a.qC2 = p->qC[2];
            a.n_phase = p->n_phase;
            a.n_atom_tab = p->n_atom * p->n_phase * kAtomTab;
            a.ax0 = p->axis[0];
            a.ax1 = p->axis[1];
            a.ax2 = p->axis[2];
        } else {
            set_strided(a, d, p->dims);
            a.in = d == 1 ? w1 : w2;
            if (d == 1)
                a.out = w2;
            else
                a.I = I;
        }
        QPD_HIP(launch_pass<T>(p->launch[d], a, d == 0, d == 2, st));
    }
    return QPD_OK;
}

// Plain 3-D transform plans, cached by (dims, precision): every field is one group
// with one chunk and the identity coefficient, so each pass is a batched 1-D FFT.
constexpr int kMaxPlainFields = 64;

struct FFTPlan {
    AxisFFT fft[3];
    Launch launch[3];
    void* dev = nullptr;
    const void* W[3];
    const Stage* stages[3];
    const void* k;       // zero wavenumbers (no derivative term is used)
    const int* index;    // 0, 1, ..., kMaxPlainFields
    const int* minus1;   // -1 repeated: identity coefficient
};

std::mutex g_fft_mutex;
std::map<std::vector<int64_t>, FFTPlan> g_fft_plans;

template <typename T>
int get_fft_plan(const int64_t dims[3], FFTPlan*& out) {
    std::vector<int64_t> key = {dims[0], dims[1], dims[2], (int64_t)sizeof(T)};
    std::lock_guard<std::mutex> lock(g_fft_mutex);
    auto it = g_fft_plans.find(key);
    if (it != g_fft_plans.end()) {
        out = &it->second;
        return QPD_OK;
    }
    FFTPlan fp;
    const int nin[3] = {1, 1, 1};
    for (int d = 0; d < 3; ++d)
        if (!make_axis(dims[d], fp.fft[d]))
            return fail(QPD_ERR_UNSUPPORTED, "unsupported transform length " + std::to_string(dims[d]));
    int rc = plan_launches(fp.fft, dims, nin, sizeof(cx<T>), fp.launch);
    if (rc) return rc;
    TableBuilder tb;
    size_t offW[3], offSt[3];
    int64_t kmax = 1;
    for (int d = 0; d < 3; ++d) {
        std::vector<double> wr(dims[d]), wi(dims[d]);
        for (int64_t m = 0; m < dims[d]; ++m) unit_root(m, dims[d], wr[m], wi[m]);
        offW[d] = tb.add_cplx<T>(wr, wi);
        offSt[d] = tb.add(fp.fft[d].st, sizeof(Stage) * kMaxStages);
        kmax = std::max(kmax, dims[d]);
    }
    const size_t offK = tb.add_real<T>(std::vector<double>(kmax, 0.0));
    std::vector<int> idx(kMaxPlainFields + 1), neg(kMaxPlainFields + 1, -1);
    for (int i = 0; i <= kMaxPlainFields; ++i) idx[i] = i;
    const size_t offI = tb.add(idx.data(), idx.size() * sizeof(int));
    const size_t offN = tb.add(neg.data(), neg.size() * sizeof(int));
    QPD_HIP(hipMalloc(&fp.dev, tb.bytes.size()));
    QPD_HIP(hipMemcpy(fp.dev, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice));
    unsigned char* base = static_cast<unsigned char*>(fp.dev);
    for (int d = 0; d < 3; ++d) {
        fp.W[d] = base + offW[d];
        fp.stages[d] = reinterpret_cast<const Stage*>(base + offSt[d]);
    }
    fp.k = base + offK;
    fp.index = reinterpret_cast<const int*>(base + offI);
    fp.minus1 = reinterpret_cast<const int*>(base + offN);
    out = &g_fft_plans.emplace(key, fp).first->second;
    return QPD_OK;
}

template <typename T>
int run_fft3d(const int64_t dims[3], void* data, int n_fields, hipStream_t st) {
    FFTPlan* fp = nullptr;
    int rc = get_fft_plan<T>(dims, fp);
    if (rc) return rc;
    cx<T>* x = static_cast<cx<T>*>(data);
    for (int d = 0; d < 3; ++d) {
        PassArgs<T> a;
        fill_common(a, fp->fft[d], fp->stages[d], fp->launch[d], dims);
        a.W = static_cast<const cx<T>*>(fp->W[d]);
        a.k = static_cast<const T*>(fp->k);
        a.chunk_src = fp->index;
        a.chunk_coef = fp->minus1;
        a.group_first = fp->index;
        a.ngroups = n_fields;
        a.nin = 1;
        a.nout = 1;
        a.in = x;
        a.out = x;
        if (d > 0) set_strided(a, d, dims);
        QPD_HIP(launch_pass<T>(fp->launch[d], a, false, false, st));
    }
    return QPD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* qpd_last_error(void) { return g_error.c_str(); }

int qpd_version(void) { return 100; }

int qpd_fft_length_supported(int64_t n) {
    std::vector<int> r;
    return factorise(n, r) && (int)r.size() <= kMaxStages && n <= 65536 ? 1 : 0;
}

int qpd_plan_create(qpd_plan** plan, const qpd_grid* grid, const qpd_crystal* c, int precision) {
    g_error.clear();
    if (!plan || !grid || !c) return fail(QPD_ERR_ARG, "null argument");
    *plan = nullptr;
    if (precision != 32 && precision != 64) return fail(QPD_ERR_ARG, "precision must be 32 or 64");
    bool seen[3] = {false, false, false};
    for (int d = 0; d < 3; ++d) {
        if (grid->dims[d] < 1) return fail(QPD_ERR_ARG, "grid dims must be >= 1");
        if (grid->axis[d] < 0 || grid->axis[d] > 2 || seen[grid->axis[d]])
            return fail(QPD_ERR_ARG, "grid axis must be a permutation of 0, 1, 2");
        seen[grid->axis[d]] = true;
        if (!(grid->h[d] > 0.0)) return fail(QPD_ERR_ARG, "grid spacing must be positive");
    }
    if (c->n_atom < 1 || c->n_atom > QPD_MAX_ATOMS) return fail(QPD_ERR_ARG, "n_atom out of range");
    if (c->n_phase < 1 || c->n_phase > QPD_MAX_PHASES) return fail(QPD_ERR_ARG, "n_phase out of range");
    if (c->n_struc != 3) return fail(QPD_ERR_UNSUPPORTED, "n_struc must be 3 (P_x, P_y, P_z)");
    if (!c->R || !c->f_re || !c->f_im || !c->su) return fail(QPD_ERR_ARG, "null crystal table");
    if (!(c->V != 0.0)) return fail(QPD_ERR_ARG, "cell volume is zero");

    if (grid->dims[0] * grid->dims[1] * grid->dims[2] >= (int64_t)1 << 31)
        return fail(QPD_ERR_UNSUPPORTED, "grids of 2^31 cells or more are not supported");
    qpd_plan* p = new qpd_plan();
    p->precision = precision;
    p->ncell = 1;
    for (int d = 0; d < 3; ++d) {
        p->dims[d] = grid->dims[d];
        p->axis[d] = grid->axis[d];
        p->ncell *= grid->dims[d];
        if (!make_axis(grid->dims[d], p->fft[d])) {
            delete p;
            return fail(QPD_ERR_UNSUPPORTED, "transform length " + std::to_string(grid->dims[d]) +
                                                 " has a prime factor above 17 or is too long");
        }
    }
    p->n_phase = c->n_phase;
    p->n_atom = c->n_atom;
    p->csize = precision == 32 ? sizeof(cx<float>) : sizeof(cx<double>);
    for (int i = 0; i < 3; ++i) p->qC[i] = c->qC[i];
    const int nin[3] = {4, 3, 2};
    int rc = plan_launches(p->fft, p->dims, nin, p->csize, p->launch, true, c->n_atom * c->n_phase * kAtomTab);
    if (rc == QPD_OK) rc = precision == 32 ? build_tables<float>(p, grid, c) : build_tables<double>(p, grid, c);
    if (rc != QPD_OK) {
        if (p->dev) (void)hipFree(p->dev);
        delete p;
        return rc;
    }
    *plan = p;
    return QPD_OK;
}

int qpd_plan_destroy(qpd_plan* plan) {
    if (!plan) return QPD_OK;
    if (plan->dev) {
        (void)hipDeviceSynchronize();
        (void)hipFree(plan->dev);
    }
    delete plan;
    return QPD_OK;
}

size_t qpd_workspace_bytes(const qpd_plan* p) {
    if (!p) return 0;
    return (size_t)(3 * p->nY + 2 * p->nZ) * (size_t)p->ncell * p->csize;
}

int qpd_launches_per_volume(const qpd_plan* p) { return p ? 3 : 0; }

int qpd_intensity_pass(qpd_plan* p, const qpd_fields* f, double* I, void* work, int pass, void* stream) {
    if (!p || !f || !I || !work) return fail(QPD_ERR_ARG, "null argument");
    if (!f->P || !f->U || !f->phase) return fail(QPD_ERR_ARG, "P, U and phase are required");
    if (pass < 0 || pass > 2) return fail(QPD_ERR_ARG, "pass must be 0, 1 or 2");
    hipStream_t st = static_cast<hipStream_t>(stream);
    return p->precision == 32 ? run_intensity<float>(p, f, I, work, st, pass)
                              : run_intensity<double>(p, f, I, work, st, pass);
}

double qpd_pass_bytes_per_cell(const qpd_plan* p, const qpd_fields* f, int pass) {
    if (!p || pass < 0 || pass > 2) return 0.0;
    const double real = p->precision == 32 ? 4.0 : 8.0, c = 2.0 * real;
    if (pass == 0) {
        double in = 3 * real + 3 * 8.0 + 1.0;  // P, U, phase
        if (f && f->region) in += real;
        if (f && f->fraction) in += p->n_phase * real;
        return in + 3.0 * p->nY * c;
    }
    if (pass == 1) return 3.0 * p->nY * c + 2.0 * p->nZ * c;
    return 2.0 * p->nZ * c + 8.0;
}

int qpd_fft3d(const int64_t dims[3], int precision, void* data, int n_fields, void* stream) {
    g_error.clear();
    if (!dims || !data) return fail(QPD_ERR_ARG, "null argument");
    if (precision != 32 && precision != 64) return fail(QPD_ERR_ARG, "precision must be 32 or 64");
    if (n_fields < 1 || n_fields > kMaxPlainFields) return fail(QPD_ERR_ARG, "n_fields must be 1..64");
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 1) return fail(QPD_ERR_ARG, "dims must be >= 1");
    if (dims[0] * dims[1] * dims[2] >= (int64_t)1 << 31)
        return fail(QPD_ERR_UNSUPPORTED, "grids of 2^31 cells or more are not supported");
    hipStream_t st = static_cast<hipStream_t>(stream);
    return precision == 32 ? run_fft3d<float>(dims, data, n_fields, st) : run_fft3d<double>(dims, data, n_fields, st);
}

int qpd_intensity(qpd_plan* p, const qpd_fields* f, double* I, void* work, void* stream) {
    if (!p || !f || !I || !work) return fail(QPD_ERR_ARG, "null argument");
    if (!f->P || !f->U || !f->phase) return fail(QPD_ERR_ARG, "P, U and phase are required");
    hipStream_t st = static_cast<hipStream_t>(stream);
    return p->precision == 32 ? run_intensity<float>(p, f, I, work, st) : run_intensity<double>(p, f, I, work, st);
}

}  // extern "C"
