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
#include "qpd_launch.hpp"
#include "qpd_ingest.hpp"
#include "qpd_project.hpp"
#include "qpd_layout.hpp"

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
// Axis factorisation: radix 8 while possible, then 4 or 2, then the odd primes up to 17
// (butterfly stages), then any larger prime factors in ascending order (direct stages,
// O(R) work per output: fft_stage_direct).
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
    for (int64_t p = 19; p * p <= m; p += 2) {
        while (m % p == 0) {
            radices.push_back((int)p);
            m /= p;
        }
    }
    if (m > 1) radices.push_back((int)m);
    return true;
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


// Whether every stage's butterflies over `lines` lines fit nthr threads at the butterflies
// a thread holds (fft_stage: mr values, mr / R butterflies).
bool stage_fits(const AxisFFT& ax, int lines, int nthr, int mr) {
    for (int s = 0; s < ax.nst; ++s) {
        const int R = ax.st[s].R;
        if (R > 17) continue;  // direct stage: any number of outputs a thread
        const int mb = std::max(1, mr / R);
        const int64_t total = (int64_t)(ax.N / R) * lines;
        if ((total + nthr - 1) / nthr > mb) return false;
    }
    return true;
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Pick (lines per field, inputs per fill, threads, EPT) for a pass over lines of length N;
// tx_cap > 0 caps the lines per field (rows of a contiguous pass, columns of a strided one).
bool choose_launch(const AxisFFT& ax, int nin, bool contiguous, size_t csize, int64_t n_lines_total,
                   int lds_soft, int lds_hard, int ept_max, Launch& L, int tx_cap = 0) {
    const int N = ax.N;
    std::vector<int> nfs;
    for (int nf : {4, 3, 2, 1})
        if (nf <= nin) nfs.push_back(nf);
    std::vector<int> per;  // lines per field candidates
    if (contiguous) {
        for (int nl = 64; nl >= 1; nl /= 2)
            if (tx_cap <= 0 || nl <= tx_cap) per.push_back(nl);
    } else {
        const int tx_pref = tx_cap > 0 ? std::min(tx_cap, csize == 8 ? 16 : 8) : (csize == 8 ? 16 : 8);
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
    // FP64 axes whose stages are all radix 2, 4 or 8 run the kRadixPow2 kernels (8 values a
    // thread) where a shape fits them, else those of radices up to 8 or of every radix (16)
    // lines of more than 4096 points (ept_max 16) run the sixteen-a-thread kernels, whose
    // stages hold kMaxRegsLong values a thread, and are compiled for every radix only
    const bool long_line = ept_max >= 16;
    bool direct = false;
    for (int s = 0; s < ax.nst; ++s) direct = direct || ax.st[s].R > 17;
    bool pow2_axis = csize == 16 && !long_line;
    bool small_axis = csize == 16 && !long_line;
    for (int s = 0; s < ax.nst; ++s) {
        pow2_axis = pow2_axis && (ax.st[s].R == 2 || ax.st[s].R == 4 || ax.st[s].R == 8);
        small_axis = small_axis && ax.st[s].R <= 8;
    }
    for (int attempt = pow2_axis ? 0 : 1; attempt < 2; ++attempt) {
        const bool pow2 = attempt == 0;
        const int mr = long_line ? kMaxRegsLong : (pow2 ? kMaxRegsPow2 : kMaxRegs);
        for (int pass = 0; pass < 2; ++pass) {
            const int cap = pass == 0 ? lds_soft : lds_hard;
            for (auto [nf, p] : order) {
                {
                    if (contiguous && p > 1 && (int64_t)p > n_lines_total) continue;
                    const int64_t elems = (int64_t)p * N;      // one field's tile
                    if (contiguous && elems > 4096 && p > 1) continue;
                    // axes with a direct stage (a prime factor above 17) hold a copy of the tile
                    const size_t lds = (size_t)nf * elems * csize * (direct ? 2 : 1);
                    if (lds > (size_t)cap) continue;
                    // threads: aim at ~ept_max/2 elements per thread, at most 512
                    int nthr = round_up((int)std::min<int64_t>(512, (elems + ept_max / 2 - 1) / (ept_max / 2)), 64);
                    nthr = std::max(64, std::min(512, nthr));
                    while (nthr < 512 && (elems + nthr - 1) / nthr > ept_max) nthr += 64;
                    if ((elems + nthr - 1) / nthr > ept_max) continue;
                    while (nthr < 512 && !stage_fits(ax, nf * p, nthr, mr)) nthr += 64;
                    if (!stage_fits(ax, nf * p, nthr, mr)) continue;
                    L.fe = nullptr;
                    L.contiguous = contiguous;
                    L.nl = p;
                    L.nf = nf;
                    L.nthr = nthr;
                    const int need = (int)((elems + nthr - 1) / nthr);
                    L.ept = long_line ? 16 : (need <= 2 ? 2 : (need <= 4 ? 4 : 8));
                    if (L.ept > ept_max) continue;
                    L.rset = pow2 ? kRadixPow2 : (small_axis ? kRadixSmall : kRadixAll);
                    L.lds = lds;
                    return true;
                }
            }
        }
    }
    return false;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device tables of a plan in one element type (FP32 or FP64): twiddles, wavenumbers and
// coefficient tables of each pass, chunk and group lists, the atom table and the
// source-group mix tables.  An FP32 plan with an FP64 tail holds both sets.
struct PlanTables {
    void* dev = nullptr;              // one allocation holding every table below
    const void* W[3] = {};
    std::map<int, const void*> Wlen;  // twiddles exp(-2 pi i m / L) of the split step lengths L
    const Stage* stages[3] = {};
    const void* k[3] = {};
    const void* coef[3] = {};
    const int* chunk_src[3] = {};
    const int* chunk_coef[3] = {};
    const int* group_first[3] = {};
    const void* atomtab = nullptr;
    const void* mixc = nullptr;       // grouped plans: X_{Y,S}(k0) tables
    const int* pair_first = nullptr;
    const int* pair_y = nullptr;
};

struct qpd_plan {
    int precision;
    int64_t dims[3];
    int axis[3];
    int64_t ncell;
    int n_phase;
    int n_atom;
    // pass d transforms memory dim pd[d]: pass 0 the contiguous dims[0], passes 1 and 2 the
    // two strided dims, by default the longer one last (its transform spreads the rounding
    // of the stored pass-1 spectra over the most points)
    int pd[3] = {0, 1, 2};
    // FP32 plans whose transforms run in FP64 on FP32 arrays (P, weights, the spectra between
    // passes): every pass's kernels are the FP64 ones with FP32 I/O
    int mixed = 0;
    // FP32 plans computed and stored in FP64: P, fractions and region are widened into the
    // workspace first (grids too small for FP32 storage between passes)
    int wide = 0;
    double est_dlog10 = 0.0;  // the error estimate the arithmetic was chosen from
    int nY, nZ;  // classes after pass 0 and pass 1
    // atom classes: y-class = equal (R[ax1], R[ax2]), z-class = equal R[ax2], with ax1, ax2
    // the physical axes of passes 1 and 2
    std::vector<int> ycls, zcls;
    std::vector<std::pair<double, double>> ykeys;
    std::vector<double> zkeys;
    // source groups: atoms with equal su (all phases) and proportional form factors
    int nS = 0;
    std::vector<int> sgrp, srep;             // group of each atom, representative of each group
    std::vector<std::complex<double>> phi;   // fE_a = phi_a fE_rep(group)
    int grouped = 0;                         // 1: pass 0 transforms groups, pass 1 mixes them
    int n1fields() const { return 3 * (grouped ? nS : nY); }
    size_t csz[3];                    // complex element bytes each pass computes in
    size_t ssz = 8;                   // complex element bytes of the spectra stored between passes
    AxisFFT fft[3];                   // transform of each pass
    PassPlan pass[3];                 // the schedule each pass runs (one of cand)
    std::vector<PassPlan> cand[3];    // schedules each pass can run, the default first
    int sel[3] = {0, 0, 0};
    int ngroups[3];
    PlanTables tab[2];                // [0] FP32, [1] FP64
    const PlanTables& tables(int d) const { return tab[csz[d] == 16 ? 1 : 0]; }
    double qC[3];
    size_t w1_bytes() const { return (size_t)n1fields() * (size_t)ncell * ssz; }
    size_t w2_bytes() const { return (size_t)2 * nZ * (size_t)ncell * ssz; }
    // a pass-2 FP64 split's intermediate (2 nZ complex FP64 fields) starts at pass 0's outputs,
    // dead by then, which the workspace places after pass 1's: the bytes it needs beyond them
    int split64 = 0;
    size_t split64_bytes() const {
        const size_t need = split64 ? (size_t)2 * nZ * (size_t)ncell * sizeof(cx<double>) : 0;
        return need > w1_bytes() ? need - w1_bytes() : 0;
    }
    // the widened inputs of a wide plan: P (3 n), fractions (n_phase n), region (n), doubles
    size_t wide_bytes() const { return wide ? (size_t)(4 + n_phase) * (size_t)ncell * sizeof(double) : 0; }
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

// Atom classes and source groups of a plan.
void classify(qpd_plan* p, const qpd_grid* g, const qpd_crystal* c) {
    const int ax1 = g->axis[p->pd[1]], ax2 = g->axis[p->pd[2]];
    const int na = c->n_atom, np = c->n_phase;
    p->ycls.assign(na, 0);
    p->ykeys.clear();
    p->zkeys.clear();
    for (int a = 0; a < na; ++a) {
        std::pair<double, double> key(c->R[3 * a + ax1], c->R[3 * a + ax2]);
        auto it = std::find(p->ykeys.begin(), p->ykeys.end(), key);
        if (it == p->ykeys.end()) {
            p->ykeys.push_back(key);
            p->ycls[a] = (int)p->ykeys.size() - 1;
        } else {
            p->ycls[a] = (int)(it - p->ykeys.begin());
        }
    }
    p->zcls.assign(p->ykeys.size(), 0);
    for (size_t y = 0; y < p->ykeys.size(); ++y) {
        auto it = std::find(p->zkeys.begin(), p->zkeys.end(), p->ykeys[y].second);
        if (it == p->zkeys.end()) {
            p->zkeys.push_back(p->ykeys[y].second);
            p->zcls[y] = (int)p->zkeys.size() - 1;
        } else {
            p->zcls[y] = (int)(it - p->zkeys.begin());
        }
    }
    p->nY = (int)p->ykeys.size();
    p->nZ = (int)p->zkeys.size();
    // source groups: same su in every phase and f_pa = phi_a f_p,rep for every phase
    auto su = [&](int d, int k, int ph, int a) { return c->su[((d * c->n_struc + k) * np + ph) * na + a]; };
    auto f = [&](int ph, int a) { return std::complex<double>(c->f_re[ph * na + a], c->f_im[ph * na + a]); };
    p->sgrp.assign(na, -1);
    p->srep.clear();
    p->phi.assign(na, 0.0);
    for (int a = 0; a < na; ++a) {
        for (int s = 0; s < (int)p->srep.size() && p->sgrp[a] < 0; ++s) {
            const int r = p->srep[s];
            bool same = true;
            for (int ph = 0; ph < np && same; ++ph)
                for (int d = 0; d < 3 && same; ++d)
                    for (int k = 0; k < 3 && same; ++k) same = su(d, k, ph, a) == su(d, k, ph, r);
            if (!same || f(0, r) == 0.0) continue;
            const std::complex<double> ratio = f(0, a) / f(0, r);
            for (int ph = 1; ph < np && same; ++ph)
                same = std::abs(f(ph, a) - ratio * f(ph, r)) <= 1e-15 * std::abs(f(ph, a));
            if (!same) continue;
            p->sgrp[a] = s;
            p->phi[a] = ratio;
        }
        if (p->sgrp[a] < 0) {
            p->sgrp[a] = (int)p->srep.size();
            p->srep.push_back(a);
            p->phi[a] = 1.0;
        }
    }
    p->nS = (int)p->srep.size();
}

template <typename T>
int build_tables(qpd_plan* p, const qpd_grid* g, const qpd_crystal* c, PlanTables& t) {
    TableBuilder tb;
    size_t offW[3], offK[3], offC[3], offSrc[3], offCoef[3], offGrp[3], offSt[3], offAtom;
    std::map<int, size_t> offWs;
    const int64_t n = p->ncell;
    const double h_of[3] = {g->h[0], g->h[1], g->h[2]};

    const int ax0 = g->axis[0];
    const std::vector<int>& ycls = p->ycls;
    const std::vector<int>& zcls = p->zcls;
    const auto& ykeys = p->ykeys;
    const auto& zkeys = p->zkeys;
    size_t offMix = 0, offPairF = 0, offPairY = 0;

    for (int d = 0; d < 3; ++d) {
        const int64_t N = g->dims[p->pd[d]];
        const double h = h_of[g->axis[p->pd[d]]];
        std::vector<double> wr(N), wi(N), kk(N);
        for (int64_t m = 0; m < N; ++m) {
            unit_root(m, N, wr[m], wi[m]);
            kk[m] = wavenumber(m, N, h);
        }
        offW[d] = tb.add_cplx<T>(wr, wi);
        offK[d] = tb.add_real<T>(kk);
        for (const PassPlan& c : p->cand[d]) {
            if (c.nlaunch != 2) continue;
            for (int L : {c.N1, c.N2}) {
                if (offWs.count(L)) continue;
                std::vector<double> sr(L), si(L);
                for (int m = 0; m < L; ++m) unit_root(m, L, sr[m], si[m]);
                offWs[L] = tb.add_cplx<T>(sr, si);
            }
        }
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
        if (d == 0 && p->grouped) {
            // pass 0 over source groups: group = source group, one chunk (its representative
            // atom) with the identity coefficient; the atom coefficients move to pass 1:
            // X_{Y,S}(k) = sum over a in Y and S of phi_a exp(-i G.R_a) / (V n) exp(-i k R_a,ax0) s(k)
            for (int sg = 0; sg < p->nS; ++sg) {
                first.push_back((int)src.size());
                src.push_back(p->srep[sg]);
                coef.push_back(-1);
            }
            first.push_back((int)src.size());
            cr.assign(1, 0.0);
            ci.assign(1, 0.0);
            std::vector<double> xr((size_t)p->nY * p->nS * N, 0.0), xi((size_t)p->nY * p->nS * N, 0.0);
            for (int a = 0; a < c->n_atom; ++a) {
                const double gr = c->G[0] * c->R[3 * a] + c->G[1] * c->R[3 * a + 1] + c->G[2] * c->R[3 * a + 2];
                const std::complex<double> scale = p->phi[a] * std::polar(1.0, -gr) / (c->V * (double)n);
                const size_t base = ((size_t)ycls[a] * p->nS + p->sgrp[a]) * N;
                for (int64_t m = 0; m < N; ++m) {
                    const std::complex<double> v =
                        scale * std::polar(1.0, -kk[m] * c->R[3 * a + ax0]) * shape_factor(kk[m], h);
                    xr[base + m] += v.real();
                    xi[base + m] += v.imag();
                }
            }
            offMix = tb.add_cplx<T>(xr, xi);
            // pairs (group S, z-class z) -> y-classes Y in z with atoms of S
            std::vector<int> pf, py;
            for (int sg = 0; sg < p->nS; ++sg)
                for (int z = 0; z < p->nZ; ++z) {
                    pf.push_back((int)py.size());
                    for (int y = 0; y < p->nY; ++y) {
                        if (zcls[y] != z) continue;
                        bool any = false;
                        for (int a = 0; a < c->n_atom; ++a) any = any || (ycls[a] == y && p->sgrp[a] == sg);
                        if (any) py.push_back(y);
                    }
                }
            pf.push_back((int)py.size());
            if (py.empty()) py.push_back(0);
            offPairF = tb.add(pf.data(), pf.size() * sizeof(int));
            offPairY = tb.add(py.data(), py.size() * sizeof(int));
        } else if (d == 0) {
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

    QPD_HIP(hipMalloc(&t.dev, tb.bytes.size()));
    QPD_HIP(hipMemcpy(t.dev, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice));
    unsigned char* base = static_cast<unsigned char*>(t.dev);
    for (int d = 0; d < 3; ++d) {
        t.W[d] = base + offW[d];
        t.stages[d] = reinterpret_cast<const Stage*>(base + offSt[d]);
        t.k[d] = base + offK[d];
        t.coef[d] = base + offC[d];
        t.chunk_src[d] = reinterpret_cast<const int*>(base + offSrc[d]);
        t.chunk_coef[d] = reinterpret_cast<const int*>(base + offCoef[d]);
        t.group_first[d] = reinterpret_cast<const int*>(base + offGrp[d]);
    }
    t.atomtab = base + offAtom;
    for (auto& kv : offWs) t.Wlen[kv.first] = base + kv.second;
    if (p->grouped) {
        t.mixc = base + offMix;
        t.pair_first = reinterpret_cast<const int*>(base + offPairF);
        t.pair_y = reinterpret_cast<const int*>(base + offPairY);
    }
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
    // unsplit addressing: element idx of a line at row idx, frequency index idx
    a.nq = 1;
    a.divQ = make_fastdiv(1);
    a.in_step = a.out_step = a.c_step = 1;
    a.Nc = ax.N;
}

// ---------------------------------------------------------------------------
// Compile-time-length kernels (FP32), looked up by (length, role).
constexpr int kMaxAtomTab = 1024;  // atom-table entries the source pass keeps in LDS

// Private (scratch) bytes per lane of a kernel, 0 when unknown.
// Static LDS of a kernel (its __shared__ arrays: the staging buffers of the staged-input
// kernels), which the device adds to the dynamic LDS of a launch.
size_t kernel_static_lds(const void* kernel) {
    if (!kernel) return 0;
    static std::mutex mu;
    static std::map<const void*, size_t> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(kernel);
    if (it != cache.end()) return it->second;
    hipFuncAttributes attr{};
    const size_t bytes = hipFuncGetAttributes(&attr, kernel) == hipSuccess ? attr.sharedSizeBytes : 0;
    cache[kernel] = bytes;
    return bytes;
}

size_t kernel_scratch_bytes(const void* kernel) {
    if (!kernel) return 0;
    static std::mutex mu;
    static std::map<const void*, size_t> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(kernel);
    if (it != cache.end()) return it->second;
    hipFuncAttributes attr{};
    const size_t bytes = hipFuncGetAttributes(&attr, kernel) == hipSuccess ? attr.localSizeBytes : 0;
    cache[kernel] = bytes;
    return bytes;
}

// The strided fast kernels address a row as a wave-uniform base plus a 32-bit lane byte
// offset of up to (64/G - 1) step S + G elements (rows step S apart, G columns): a
// schedule whose offsets could reach 4 GiB is not offered.
bool lane_offsets_fit(int G, int64_t step, int64_t S, size_t csize) {
    const int64_t lanes = 64 / std::max(G, 1) - 1;
    return (double)(lanes * step * S + G) * (double)csize < 4294967296.0;
}

void set_fast(Launch& L, const FastEntry* fe, size_t lds) {
    L = Launch();
    L.fe = fe;
    L.contiguous = fe->role <= kRoleContig || fe->role == kRoleSourceW;
    L.nl = fe->G;
    L.nf = (fe->role == kRoleSource || fe->role == kRoleSourceW) ? 4 : 1;
    L.nthr = fe->nthr;
    L.lds = lds;
}

// Kernel instance for a pass: contiguous passes compute the source (3 outputs) or
// transform fields (1 output); strided passes have 2 outputs, or 1 when they
// write the intensity or run a plain transform.
// The generic kernels of one pass; the kRadixPow2 and kRadixSmall instantiations only for
// FP64 (the FP32 forms of every radix spill little).  Elements a thread: the smallest instance holding L.ept; 16 only for lines
// of more than 4096 points (up to 8192: 512 threads of 16), which spill but run.
template <typename T>
using GenericKernel = void (*)(PassArgs<T>);

template <typename T>
hipError_t launch_ept(const Launch& L, const PassArgs<T>& a, hipStream_t st, GenericKernel<T> k2, GenericKernel<T> k4,
                      GenericKernel<T> k8, GenericKernel<T> k16) {
    GenericKernel<T> k = L.ept <= 2 ? k2 : L.ept <= 4 ? k4 : L.ept <= 8 ? k8 : L.ept <= 16 ? k16 : nullptr;
    if (!k) return hipErrorInvalidValue;
    return launch_one(k, L, st, a);
}

template <typename T, int RS, bool IO32, int NOUT, bool SOURCE>
hipError_t launch_contig(const Launch& L, const PassArgs<T>& a, hipStream_t st) {
    GenericKernel<T> k16 = nullptr;
    if constexpr (RS == kRadixAll) k16 = pass_contig<T, 16, NOUT, SOURCE, RS, IO32>;
    // two elements a thread where they suffice: the FP64 source kernel holds each element's
    // cell inputs and accumulators across the groups; the plain transform starts at four
    return launch_ept<T>(L, a, st,
                         SOURCE ? pass_contig<T, 2, NOUT, SOURCE, RS, IO32> : pass_contig<T, 4, NOUT, SOURCE, RS, IO32>,
                         pass_contig<T, 4, NOUT, SOURCE, RS, IO32>, pass_contig<T, 8, NOUT, SOURCE, RS, IO32>, k16);
}

template <typename T, int RS, bool IO32, int NOUT, bool INTENSITY>
hipError_t launch_strided(const Launch& L, const PassArgs<T>& a, hipStream_t st) {
    GenericKernel<T> k16 = nullptr;
    if constexpr (RS == kRadixAll) k16 = pass_strided<T, 16, NOUT, INTENSITY, RS, IO32>;
    // four elements a thread where they suffice: fewer registers (FP64 kernels spill)
    return launch_ept<T>(L, a, st, pass_strided<T, 4, NOUT, INTENSITY, RS, IO32>,
                         pass_strided<T, 4, NOUT, INTENSITY, RS, IO32>, pass_strided<T, 8, NOUT, INTENSITY, RS, IO32>,
                         k16);
}

template <typename T, int RS, bool IO32, int NZ>
hipError_t launch_mix(const Launch& L, const PassArgs<T>& a, hipStream_t st) {
    GenericKernel<T> k16 = nullptr;
    if constexpr (RS == kRadixAll) k16 = pass_strided_mix<T, 16, NZ, RS, IO32>;
    return launch_ept<T>(L, a, st, pass_strided_mix<T, 4, NZ, RS, IO32>, pass_strided_mix<T, 4, NZ, RS, IO32>,
                         pass_strided_mix<T, 8, NZ, RS, IO32>, k16);
}

template <typename T, int RS, bool IO32>
hipError_t launch_generic_io(const Launch& L, const PassArgs<T>& a, bool source, bool intensity, hipStream_t st) {
    if (L.contiguous) {
        // FP64 on FP32 arrays: only the source pass (plain transforms never read FP32 spectra)
        if (source) return a.nout != 3 ? hipErrorInvalidValue : launch_contig<T, RS, IO32, 3, true>(L, a, st);
        if (IO32 || a.nout != 1) return hipErrorInvalidValue;
        return launch_contig<T, RS, IO32, 1, false>(L, a, st);
    }
    if (a.pair_first) {  // pass 1 over source groups: one or two z-classes
        if (a.ngroups == 1) return launch_mix<T, RS, IO32, 1>(L, a, st);
        if (a.ngroups == 2) return launch_mix<T, RS, IO32, 2>(L, a, st);
        return hipErrorInvalidValue;
    }
    if (intensity) return a.nout != 1 ? hipErrorInvalidValue : launch_strided<T, RS, IO32, 1, true>(L, a, st);
    if (a.nout == 2) return launch_strided<T, RS, IO32, 2, false>(L, a, st);
    if (IO32 || a.nout != 1) return hipErrorInvalidValue;
    return launch_strided<T, RS, IO32, 1, false>(L, a, st);
}

template <typename T, int RS>
hipError_t launch_generic(const Launch& L, const PassArgs<T>& a, bool source, bool intensity, hipStream_t st) {
    if constexpr (std::is_same<T, double>::value) {
        // an FP32 plan computing in FP64: FP32 P, weights and spectra
        if (L.io32) return launch_generic_io<T, RS, true>(L, a, source, intensity, st);
    }
    return launch_generic_io<T, RS, false>(L, a, source, intensity, st);
}

template <typename T>
hipError_t launch_pass(const Launch& L, const PassArgs<T>& a, bool source, bool intensity, hipStream_t st) {
    if (L.fe) {
        if constexpr (std::is_same<T, float>::value)
            return L.fe->launch ? L.fe->launch(L, a, st) : hipErrorInvalidValue;
        else
            return L.fe->launch64 ? L.fe->launch64(L, a, st) : hipErrorInvalidValue;
    }
    if constexpr (std::is_same<T, double>::value) {
        if (L.rset == kRadixPow2) return launch_generic<T, kRadixPow2>(L, a, source, intensity, st);
        if (L.rset == kRadixSmall) return launch_generic<T, kRadixSmall>(L, a, source, intensity, st);
    }
    return launch_generic<T, kRadixAll>(L, a, source, intensity, st);
}

// Schedules of pass d with nin inputs per chunk, the default first: the registered
// compile-time kernels of the length and role (FP32), then, for strided lengths, every
// split N1 x N2 into registered step lengths (most even first; only where allow_split),
// then the generic kernel (for strided passes at two tile widths).  A split is the default
// only where the length has no kernel of its own.
// Pass d runs along memory dim md (pass 0 along dims[0]); io32: an FP64 pass on FP32 arrays
// (an FP32 plan computing in FP64).
int pass_candidates(const AxisFFT& ax, int d, int md, const int64_t dims[3], int nin, size_t csize, bool io32,
                    bool intensity_last, int atomtab_entries, bool allow_split, std::vector<PassPlan>& out,
                    int mix_role = -1, int nsrc = 0, int ncoef = 0) {
    out.clear();
    const int hard = max_lds_bytes();
    const int soft = std::min(hard, 65536);
    const bool contig = d == 0;
    const int64_t N = dims[md];
    const int64_t lines = contig ? dims[1] * dims[2] : 0;
    const int64_t n_other = md == 1 ? dims[2] : dims[1];
    const int64_t row_stride = md == 1 ? dims[0] : dims[0] * dims[1];
    // compile-time kernels: every role in FP32; the strided roles and the block-per-row
    // source pass in FP64
    if (fast_enabled()) {
        int role = -1;
        if (contig)
            role = (nin == 4 && atomtab_entries <= kMaxAtomTab) ? kRoleSource : (nin == 1 ? kRoleContig : -1);
        else if (nin == 3)
            role = (d == 1 && mix_role >= 0) ? mix_role : kRoleStrided;
        else if (nin == 2 && intensity_last && d == 2)
            role = kRoleIntensity;
        else if (nin == 1)
            role = kRoleStrided1;
        if (role >= 0) {
            // source-group plans (mix_role >= 0) run the one-wave-per-row source pass first,
            // unless its registers spill more than 32 bytes a lane, in which case it follows
            // the block-per-row kernels
            const bool wave = role == kRoleSource && mix_role >= 0;
            const int roles[3] = {wave ? kRoleSourceW : -1, role, wave ? kRoleSourceW : -1};
            for (int pos = 0; pos < 3; ++pos) {
                const int rl = roles[pos];
                if (rl < 0) continue;
                for (const FastEntry& e : fast_registry()) {
                    if (e.N != N || e.role != rl || (size_t)e.real_bytes * 2 != csize || e.io32 != io32) continue;
                    if (e.nsrc != 0 && e.nsrc != nsrc) continue;
                    if (rl == kRoleSourceW && (pos == 0) != (kernel_scratch_bytes(e.kernel) <= 32)) continue;
                    const bool src = rl == kRoleSource || rl == kRoleSourceW;
                    // the held pass over source groups stages the pass's coefficient tables and
                    // the tile's mix factors (ncoef (N + nsrc G) values) in LDS
                    const size_t lds = e.lds + (src ? csize / 2 * (size_t)atomtab_entries : 0) +
                                       (e.nsrc || e.tab_lds
                                            ? csize * (size_t)ncoef * (size_t)(N + (e.nsrc ? e.nsrc : nsrc) * e.G)
                                            : 0);
                    if ((int)(lds + (e.lds_stage ? kernel_static_lds(e.kernel) : 0)) > hard) continue;
                    if (!contig && !lane_offsets_fit(e.G, 1, row_stride, csize)) continue;
                    // sixteen-byte staged loads: 16-byte aligned column pairs, rows and fields
                    if (e.lds_stage && (dims[0] % 2 != 0 || row_stride % 2 != 0 ||
                                        (dims[0] * dims[1] * dims[2]) % 2 != 0))
                        continue;
                    PassPlan pp;
                    set_fast(pp.L[0], &e, lds);
                    pp.L[0].blocks = contig ? (lines + e.G - 1) / e.G : (dims[0] + e.G - 1) / e.G * n_other;
                    if (pp.L[0].blocks <= 0x7fffffffLL) out.push_back(pp);
                }
            }
            const bool splittable =
                role == kRoleStrided || role == kRoleIntensity || role == kRoleMix1 || role == kRoleMix2;
            // FP64 transforms on FP32 spectra split only the intensity pass of a long axis: its
            // first step writes an FP64 intermediate (FP32 storage there would round partial
            // transforms of the line) into the workspace (qpd_plan::split64_bytes)
            const bool split64 = csize == 16 && io32 && role == kRoleIntensity && N > 1024;
            if (!contig && allow_split && splittable && (csize == 8 || split64)) {
                std::vector<std::pair<double, PassPlan>> splits;
                for (int64_t n1 = 2; n1 <= N / 2; ++n1) {
                    if (N % n1) continue;
                    const int64_t n2 = N / n1;
                    for (const FastEntry& ea : fast_registry()) {
                        if (ea.N != n2 || ea.role != kRoleSplit ||
                            (int)(ea.lds + (ea.lds_stage ? kernel_static_lds(ea.kernel) : 0)) > hard)
                            continue;
                        if ((size_t)ea.real_bytes * 2 != csize || ea.io32 != split64) continue;
                        if (ea.lds_stage && (dims[0] % 2 != 0 || row_stride % 2 != 0 ||
                                             (dims[0] * dims[1] * dims[2]) % 2 != 0))
                            continue;
                        if (!lane_offsets_fit(ea.G, n1, row_stride, csize)) continue;
                        const int role_b = role == kRoleStrided     ? kRoleStridedB
                                           : role == kRoleIntensity ? kRoleIntensityB
                                           : role == kRoleMix1      ? kRoleMix1B
                                                                    : kRoleMix2B;
                        for (const FastEntry& eb : fast_registry()) {
                            if (eb.N != n1 || eb.role != role_b || (int)eb.lds > hard) continue;
                            if ((size_t)eb.real_bytes * 2 != csize || eb.io32) continue;
                            // tables of the pass over source groups staged in LDS (step 2: N1 points)
                            const size_t ldsb = eb.lds + (eb.tab_lds ? csize * (size_t)ncoef * (size_t)(n1 + nsrc * eb.G) : 0);
                            if ((int)(ldsb + (eb.lds_stage ? kernel_static_lds(eb.kernel) : 0)) > hard) continue;
                            if (eb.lds_stage && (dims[0] % 2 != 0 || row_stride % 2 != 0 ||
                                                 (dims[0] * dims[1] * dims[2]) % 2 != 0))
                                continue;
                            if (!lane_offsets_fit(eb.G, n2, row_stride, csize)) continue;
                            PassPlan pp;
                            pp.nlaunch = 2;
                            pp.N1 = (int)n1;
                            pp.N2 = (int)n2;
                            set_fast(pp.L[0], &ea, ea.lds);
                            set_fast(pp.L[1], &eb, ldsb);
                            pp.L[0].blocks = (dims[0] + ea.G - 1) / ea.G * n1 * n_other;
                            pp.L[1].blocks = (dims[0] + eb.G - 1) / eb.G * n2 * n_other;
                            if (pp.L[0].blocks > 0x7fffffffLL || pp.L[1].blocks > 0x7fffffffLL) continue;
                            const double score = std::fabs(std::log((double)n1 / (double)n2)) - 1e-6 * (double)n1;
                            splits.push_back({score, pp});
                        }
                    }
                }
                std::stable_sort(splits.begin(), splits.end(),
                                 [](const auto& x, const auto& y) { return x.first < y.first; });
                for (auto& sp : splits) out.push_back(sp.second);
            }
        }
    }
    PassPlan gp;
    Launch& L = gp.L[0];
    // the source pass holds at most four cells a thread, or eight where a row of more than
    // 2048 cells leaves no other shape (512 threads a block)
    int ept_max = contig ? (nin > 1 ? 4 : 8) : 8;
    bool fits = choose_launch(ax, nin, contig, csize, lines, soft, hard, ept_max, L);
    if (!fits && contig && nin > 1) {
        ept_max = 8;
        fits = choose_launch(ax, nin, contig, csize, lines, soft, hard, ept_max, L);
    }
    if (!fits) {  // lines of more than 4096 points: sixteen a thread
        ept_max = 16;
        fits = choose_launch(ax, nin, contig, csize, lines, soft, hard, ept_max, L);
    }
    if (fits) {
        L.io32 = io32 ? 1 : 0;
        L.mix = mix_role >= 0 && d == 1 ? 1 : 0;
        L.blocks = contig ? (lines + L.nl - 1) / L.nl : (dims[0] + L.nl - 1) / L.nl * n_other;
        const bool pushed = L.blocks <= 0x7fffffffLL;
        if (pushed) out.push_back(gp);
        // the half-width tile as a second generic schedule, for the autotuner to weigh:
        // strided passes get fewer elements and registers a thread and narrower rows,
        // contiguous ones half the rows (half the LDS and threads: more blocks per CU)
        if (L.nl > 1) {
            PassPlan gp2;
            Launch& L2 = gp2.L[0];
            if (choose_launch(ax, nin, contig, csize, lines, soft, hard, ept_max, L2, L.nl / 2)) {
                L2.io32 = io32 ? 1 : 0;
                L2.mix = mix_role >= 0 && d == 1 ? 1 : 0;
                L2.blocks = contig ? (lines + L2.nl - 1) / L2.nl : (dims[0] + L2.nl - 1) / L2.nl * n_other;
                // FP64 source passes default to the half tile: at about 165 VGPRs a 512-thread
                // block of 64 KB runs alone on its CU (3.55 against 2.60 ms at 512x512x104)
                const bool first = pushed && contig && csize == 16 && atomtab_entries > 0;
                if (L2.blocks <= 0x7fffffffLL) out.insert(first ? out.end() - 1 : out.end(), gp2);
            }
        }
    }
    if (out.empty())
        return fail(QPD_ERR_UNSUPPORTED, "no launch shape fits dim " + std::to_string(d) + " of length " + std::to_string(N));
    return QPD_OK;
}

// Default schedules of the three passes (plain transforms never split: the second step
// of a split writes to another buffer).
int plan_launches(const AxisFFT fft[3], const int64_t dims[3], const int nin[3], size_t csize, PassPlan out[3]) {
    for (int d = 0; d < 3; ++d) {
        std::vector<PassPlan> c;
        int rc = pass_candidates(fft[d], d, d, dims, nin[d], csize, false, false, 0, false, c);
        if (rc) return rc;
        out[d] = c[0];
    }
    return QPD_OK;
}

template <typename T>
void set_strided(PassArgs<T>& a, int d, const int64_t dims[3]) {
    a.dim = d;
    a.ntiles0 = (uint32_t)((dims[0] + a.nl - 1) / a.nl);
    if (d == 1) {
        a.S = (uint32_t)dims[0];
        a.OS = (uint32_t)(dims[0] * dims[1]);
    } else {
        a.S = (uint32_t)(dims[0] * dims[1]);
        a.OS = (uint32_t)dims[0];
    }
}

// Enqueue pass d of a volume in element type T (the pass's own precision).  The workspace
// holds pass 1's outputs (w2) followed by pass 0's (w1), both in the storage precision, then
// the FP64 intermediate of a split intensity pass beyond w1 (which it overwrites: pass 0's
// outputs are dead once pass 1 has run), or the widened inputs of a wide plan.
template <typename T>
int run_pass(qpd_plan* p, const qpd_fields* f, double* I, void* work, int d, hipStream_t st) {
    const PlanTables& tb = p->tables(d);
    unsigned char* w2 = static_cast<unsigned char*>(work);
    unsigned char* w1 = w2 + p->w2_bytes();
    const int md = p->pd[d];
    const PassPlan& pp = p->pass[d];
    const Launch& L = pp.L[pp.nlaunch - 1];  // the launch that accumulates
    PassArgs<T> a;
    fill_common(a, p->fft[d], tb.stages[d], L, p->dims);
    a.W = static_cast<const cx<T>*>(tb.W[d]);
    a.coef = static_cast<const cx<T>*>(tb.coef[d]);
    a.k = static_cast<const T*>(tb.k[d]);
    a.chunk_src = tb.chunk_src[d];
    a.chunk_coef = tb.chunk_coef[d];
    a.group_first = tb.group_first[d];
    a.ngroups = p->ngroups[d];
    a.nin = 4 - d;
    a.nout = 3 - d;
    if (d == 0) {
        a.out = reinterpret_cast<cx<T>*>(w1);
        a.P = static_cast<const T*>(f->P);
        a.U = f->U;
        a.phase = f->phase;
        a.fraction = static_cast<const T*>(f->fraction);
        a.region = static_cast<const T*>(f->region);
        a.atomtab = static_cast<const T*>(tb.atomtab);
        a.qC0 = p->qC[0];
        a.qC1 = p->qC[1];
        a.qC2 = p->qC[2];
        a.n_phase = p->n_phase;
        a.n_atom_tab = p->n_atom * p->n_phase * kAtomTab;
        // the fields pass 0 writes: [E - i k0 U_ax0, U_ax1, U_ax2] with ax1, ax2 the physical
        // axes that passes 1 and 2 transform
        a.ax0 = p->axis[0];
        a.ax1 = p->axis[p->pd[1]];
        a.ax2 = p->axis[p->pd[2]];
    } else {
        set_strided(a, md, p->dims);
        // the spectra of a plan computing in FP64 on FP32 arrays are passed through pointers of
        // the compute type; its kernels (io32) reinterpret them
        a.in = reinterpret_cast<const cx<T>*>(d == 1 ? w1 : w2);
        if (d == 1)
            a.out = reinterpret_cast<cx<T>*>(w2);
        else
            a.I = I;
        if (d == 1 && p->grouped) {
            a.nsrc = p->nS;
            a.ncoef = p->nY;
            a.pair_first = tb.pair_first;
            a.pair_y = tb.pair_y;
            a.mixc = static_cast<const cx<T>*>(tb.mixc);
        }
    }
    if (pp.nlaunch == 2) {
        // step 1: N2-point sub-transforms of every input field, twiddled, in place; a pass
        // computing in FP64 on FP32 spectra (the intensity pass only) writes them to the FP64
        // intermediate at w1 instead, and step 2 reads that
        const bool split64 = p->mixed && p->csz[d] == 16;
        if (split64 && (d != 2 || !p->split64))
            return fail(QPD_ERR_UNSUPPORTED, "FP64 split of a pass other than the intensity pass");
        const uint32_t N1 = (uint32_t)pp.N1, N2 = (uint32_t)pp.N2;
        PassArgs<T> s1 = a;
        s1.N = (int)N2;
        s1.nl = pp.L[0].nl;
        set_strided(s1, md, p->dims);
        s1.W = static_cast<const cx<T>*>(tb.Wlen.at(pp.N2));
        s1.Wsplit = static_cast<const cx<T>*>(tb.W[d]);
        s1.nq = N1;
        s1.divQ = make_fastdiv(N1);
        s1.in_qmul = s1.out_qmul = 1;
        s1.in_step = s1.out_step = N1;
        s1.out = split64 ? reinterpret_cast<cx<T>*>(w1) : const_cast<cx<T>*>(a.in);
        s1.nfields = (d == 1 ? p->n1fields() : 2 * p->nZ);
        QPD_HIP(launch_pass<T>(pp.L[0], s1, false, false, st));
        if (split64) a.in = reinterpret_cast<const cx<T>*>(w1);
        // step 2: N1-point transforms over rows N1 q + y1 into rows q + N2 k1, accumulated
        a.N = (int)N1;
        a.W = static_cast<const cx<T>*>(tb.Wlen.at(pp.N1));
        a.nq = N2;
        a.divQ = make_fastdiv(N2);
        a.in_qmul = N1;
        a.in_step = 1;
        a.out_qmul = 1;
        a.out_step = N2;
        a.c_qmul = 1;
        a.c_step = N2;
    }
    QPD_HIP(launch_pass<T>(L, a, d == 0, d == 2, st));
    return QPD_OK;
}

// dst[i] = src[i] in double, i < n (grid-stride).
__global__ __launch_bounds__(256) void widen_kernel(const float* __restrict__ src, double* __restrict__ dst, size_t n) {
    const size_t step = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) dst[i] = (double)src[i];
}

hipError_t widen(const void* src, double* dst, size_t n, hipStream_t st) {
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(widen_kernel, dim3((unsigned)blocks), dim3(256), 0, st, static_cast<const float*>(src), dst, n);
    return hipGetLastError();
}

int run_intensity(qpd_plan* p, const qpd_fields* f, double* I, void* work, hipStream_t st, int only = -1) {
    for (int d = 0; d < 3; ++d) {
        if (only >= 0 && d != only) continue;
        int rc;
        if (d == 0 && p->wide) {
            // FP32 inputs of a plan computed in FP64: widened into the end of the workspace
            const size_t n = (size_t)p->ncell;
            double* Pd = reinterpret_cast<double*>(static_cast<unsigned char*>(work) + p->w1_bytes() + p->w2_bytes());
            double* fr = Pd + 3 * n;
            double* rg = fr + (size_t)p->n_phase * n;
            qpd_fields wf = *f;
            QPD_HIP(widen(f->P, Pd, 3 * n, st));
            wf.P = Pd;
            if (f->fraction) {
                QPD_HIP(widen(f->fraction, fr, (size_t)p->n_phase * n, st));
                wf.fraction = fr;
            }
            if (f->region) {
                QPD_HIP(widen(f->region, rg, n, st));
                wf.region = rg;
            }
            rc = run_pass<double>(p, &wf, I, work, d, st);
        } else {
            rc = p->csz[d] == 8 ? run_pass<float>(p, f, I, work, d, st) : run_pass<double>(p, f, I, work, d, st);
        }
        if (rc != QPD_OK) return rc;
    }
    return QPD_OK;
}

// Scratch of qpd_intensity_noalloc: the runtime gives every device the code object loads on
// its own copy of this array.  Host code never passes the symbol itself to a kernel or a copy
// (on the host it names a stand-in, not a device address): each device's address is looked
// up once, with that device current, and kept in g_scratch.
__device__ __align__(256) unsigned char scratch[QPD_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch[kMaxScratchDevices] = {};

// The scratch address on the current device (no allocation: a symbol lookup on the first
// call per device, a table read after it).
int scratch_of_current_device(void** out) {
    int dev = 0;
    QPD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return fail(QPD_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        QPD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(QPD_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return QPD_OK;
}

// Plain 3-D transform plans, cached by (dims, precision): every field is one group
// with one chunk and the identity coefficient, so each pass is a batched 1-D FFT.
constexpr int kMaxPlainFields = 64;

struct FFTPlan {
    AxisFFT fft[3];
    PassPlan pass[3];
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
    int rc = plan_launches(fp.fft, dims, nin, sizeof(cx<T>), fp.pass);
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
        fill_common(a, fp->fft[d], fp->stages[d], fp->pass[d].L[0], dims);
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
        QPD_HIP(launch_pass<T>(fp->pass[d].L[0], a, false, false, st));
    }
    return QPD_OK;
}

std::string describe_launch(const Launch& L) {
    char buf[160];
    static const char* role_name[kNumRoles] = {"source", "contiguous", "strided", "intensity", "strided1",
                                               "split-step", "strided", "intensity", "groups", "groups",
                                               "groups", "groups", "source-wave"};
    if (L.fe) {
        std::snprintf(buf, sizeof(buf), "fast %d (%s) G%d %s%s%s%s%s%s%s", L.fe->N, L.fe->radices, L.fe->G,
                      role_name[L.fe->role], L.fe->nsrc ? (L.fe->nsrc == 2 ? " held2" : " held3") : "",
                      L.fe->split > 1 ? (" split" + std::to_string(L.fe->split)).c_str() : "",
                      L.fe->no_prefetch ? " np" : "", L.fe->real_bytes == 8 ? " f64" : "", L.fe->io32 ? " io32" : "",
                      L.fe->one_buffer ? " 1buf" : "");
        if (L.fe->lds_stage) std::strncat(buf, " ldsin", sizeof(buf) - std::strlen(buf) - 1);
        if (L.fe->tab_lds) std::strncat(buf, " tabl", sizeof(buf) - std::strlen(buf) - 1);
    } else
        std::snprintf(buf, sizeof(buf), "generic %s%s nl%d nf%d nthr%d ept%d%s%s", L.contiguous ? "contiguous" : "strided",
                      L.mix ? " groups" : "", L.nl, L.nf, L.nthr, L.ept,
                      L.rset == kRadixPow2 ? " pow2" : (L.rset == kRadixSmall ? " small" : ""), L.io32 ? " io32" : "");
    return buf;
}

std::string describe(const PassPlan& pp) {
    if (pp.nlaunch == 1) return describe_launch(pp.L[0]);
    return "split " + std::to_string(pp.N1 * pp.N2) + " = " + std::to_string(pp.N1) + " x " + std::to_string(pp.N2) +
           ": " + describe_launch(pp.L[0]) + " + " + describe_launch(pp.L[1]);
}

}  // namespace

// ---------------------------------------------------------------------------
namespace qpd {

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

const std::vector<FastEntry>& fast_registry() {
    static const std::vector<FastEntry> reg = [] {
        std::vector<FastEntry> v;
        register_fast_part0(v);
        register_fast_part1(v);
        register_fast_part2(v);
        register_fast_part3(v);
        register_fast_part4(v);
        register_fast_part5(v);
        register_fast_part6(v);
        register_fast_part7(v);
        return v;
    }();
    return reg;
}

}  // namespace qpd

// ---------------------------------------------------------------------------
namespace {
template <typename TP, typename TU, typename TO>
hipError_t launch_ingest(const IngestArgs& a, void* P_out, double* U_out, signed char* ph, hipStream_t st) {
    const unsigned blocks = (a.nc + 255u) / 256u;
    hipLaunchKernelGGL((ingest_columns<TP, TU, TO>), dim3(blocks), dim3(256), 0, st, a, static_cast<TO*>(P_out), U_out, ph);
    return hipGetLastError();
}

template <typename TO>
hipError_t ingest_dispatch(const IngestArgs& a, int pb, int ub, void* P_out, double* U_out, signed char* ph,
                           hipStream_t st) {
    if (pb == 32)
        return ub == 64 ? launch_ingest<float, double, TO>(a, P_out, U_out, ph, st)
                        : launch_ingest<float, float, TO>(a, P_out, U_out, ph, st);
    return ub == 64 ? launch_ingest<double, double, TO>(a, P_out, U_out, ph, st)
                    : launch_ingest<double, float, TO>(a, P_out, U_out, ph, st);
}
}  // namespace

extern "C" {

const char* qpd_last_error(void) { return g_error.c_str(); }

int qpd_version(void) { return 100; }

int qpd_fft_length_supported(int64_t n) {
    std::vector<int> r;
    return factorise(n, r) && (int)r.size() <= kMaxStages && n <= 65536 ? 1 : 0;
}

int qpd_plan_create(qpd_plan** plan, const qpd_grid* grid, const qpd_crystal* c, int precision) {
    return qpd_plan_create_opts(plan, grid, c, precision, QPD_ORDER_AUTO, QPD_ARITH_AUTO);
}

int qpd_plan_create_opts(qpd_plan** plan, const qpd_grid* grid, const qpd_crystal* c, int precision, int order,
                         int arith) {
    g_error.clear();
    if (!plan || !grid || !c) return fail(QPD_ERR_ARG, "null argument");
    *plan = nullptr;
    if (precision != 32 && precision != 64) return fail(QPD_ERR_ARG, "precision must be 32 or 64");
    if (order != QPD_ORDER_AUTO && order != QPD_ORDER_MEMORY && order != QPD_ORDER_SWAPPED)
        return fail(QPD_ERR_ARG, "order must be QPD_ORDER_AUTO, QPD_ORDER_MEMORY or QPD_ORDER_SWAPPED");
    if (arith < QPD_ARITH_AUTO || arith > QPD_ARITH_FP64)
        return fail(QPD_ERR_ARG, "arith must be one of QPD_ARITH_AUTO, _FP32, _MIXED, _FP64");
    if (precision == 64 && arith != QPD_ARITH_AUTO && arith != QPD_ARITH_FP64)
        return fail(QPD_ERR_ARG, "a precision-64 plan computes in FP64 throughout");
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
    }
    // source groups: when they are fewer than the y-classes, pass 0 transforms one
    // representative per group and pass 1 mixes them into one or two z-classes (classes of
    // equal position along the last pass's axis); QPD_CLASS_FORM=1 keeps the class form for
    // comparison runs
    const char* cf = std::getenv("QPD_CLASS_FORM");
    const bool class_form = cf && cf[0] == '1';
    auto groups_with = [&](int pd1) {
        p->pd[1] = pd1;
        p->pd[2] = 3 - pd1;
        classify(p, grid, c);
        return p->nZ <= 2 && p->nS < p->nY && !class_form;
    };
    // Arithmetic of FP32 plans.  An FP32 transform of a line that holds a strong peak leaves
    // errors of order eps |A_peak| on its other outputs (the rounding inside the butterflies
    // that carry the peak, whatever the radix order, twiddle precision or decimation), and a
    // voxel at 1e-10 of the peak has |A| = 1e-5 |A_peak|: FP32 arithmetic gives |dlog10 I| of
    // 1e-3 and more there (measured 0.9e-3 to 1.8e-3 on lines of 20 to 2000 points, NumPy
    // emulation of the FP32 passes and the GPU results), and where the lines of a volume are
    // alike these errors repeat from line to line and are not averaged down by the later
    // transforms.  So the transforms run in FP64.  The spectra between passes stay FP32: the
    // rounding of a stored value is relative to that value, so where lines repeat it follows
    // the signal, and otherwise the FP64 transforms after it spread it out: pass 1's along the
    // last axis, kLineErr / sqrt(N_last) (none when N_last = 1: the stored values are then the
    // amplitudes), and pass 0's over the strided plane, kPlaneErr / sqrt(N1 N2) (none when the
    // plane is one cell).  The constants are the worst cases measured per term (NumPy
    // emulation of FP32 storage with FP64 transforms).  A plan whose estimate exceeds kBound,
    // 0.6 of the 1e-3 target, stores its spectra in FP64 too.
    // A long strided pass may still run FP32 transforms: their errors on a line are about
    // eps sqrt(log2 N) times the rms of the line's outputs, so the worst case (a voxel at the
    // floor on the line of a single bright peak) falls as sqrt(log2 N / N); kTxErr is that
    // worst case's constant, measured on the 1024 x 1024 x 1 hexagonal-twin example (1.39e-3
    // over two 1024-point FP32 passes, the worst result of FP32 transforms in the tests).  The
    // automatic arithmetic gives FP32 transforms to pass 1, then to pass 2, while the estimate
    // with their terms stays within kBound (passes of about 3200 points and more; one of the two
    // 4000-point passes of a 4000 x 4000 x 68 grid).  Axes with a direct stage (a prime factor
    // above 17, whose outputs are long FP32 sums) keep FP64 transforms: the model is the
    // butterflies'.
    constexpr double kLineErr = 5.8e-3, kPlaneErr = 3.5e-3, kBound = 6e-4, kTxErr = 9.8e-3;
    auto spread = [](double k, double n) { return n > 1.0 ? k / std::sqrt(n) : 0.0; };
    auto estimate = [&](int pd1) {
        const double nl = (double)grid->dims[3 - pd1], n1 = (double)grid->dims[pd1];
        return std::hypot(spread(kLineErr, nl), spread(kPlaneErr, n1 * nl));
    };
    // pass order: an order that keeps the source-group form, then the order of the smaller
    // estimate (ties keep the memory order)
    bool swap = order == QPD_ORDER_SWAPPED;
    if (order == QPD_ORDER_AUTO) {
        const bool g_mem = groups_with(1), g_swap = groups_with(2);
        swap = g_mem != g_swap ? g_swap : estimate(2) < estimate(1);
    }
    p->pd[1] = swap ? 2 : 1;
    p->pd[2] = swap ? 1 : 2;
    for (int d = 0; d < 3; ++d) {
        const int64_t N = grid->dims[p->pd[d]];
        if (!make_axis(N, p->fft[d])) {
            delete p;
            return fail(QPD_ERR_UNSUPPORTED, "transform length " + std::to_string(N) +
                                                 " has a prime factor above 17 or is too long");
        }
    }
    bool fp32_pass[3] = {false, false, false};  // passes of a mixed plan that keep FP32 transforms
    if (precision == 32) {
        int a = arith;
        double est = estimate(p->pd[1]);
        const bool automatic = a == QPD_ARITH_AUTO;
        if (automatic) a = est <= kBound ? QPD_ARITH_MIXED : QPD_ARITH_FP64;
        p->mixed = a == QPD_ARITH_MIXED ? 1 : 0;
        p->wide = a == QPD_ARITH_FP64 ? 1 : 0;
        if (automatic && p->mixed) {
            for (int d = 1; d <= 2; ++d) {
                const double n = (double)grid->dims[p->pd[d]];
                const double tx = kTxErr * std::sqrt(std::log2(n) / n);
                bool butterflies = true;
                for (int s = 0; s < p->fft[d].nst; ++s) butterflies = butterflies && p->fft[d].st[s].R <= 17;
                if (n > 1.0 && butterflies && std::hypot(est, tx) <= kBound) {
                    fp32_pass[d] = true;
                    est = std::hypot(est, tx);
                }
            }
        }
        // an explicit FP32 plan has no bound: its FP32 transforms leave about 1e-3 at voxels
        // 1e-10 of the peak on lines that carry a strong one
        p->est_dlog10 = p->wide ? 0.0 : p->mixed ? est : -1.0;
    }
    const bool f64 = precision == 64 || p->wide || p->mixed;
    for (int d = 0; d < 3; ++d) p->csz[d] = f64 && !fp32_pass[d] ? sizeof(cx<double>) : sizeof(cx<float>);
    p->ssz = (precision == 64 || p->wide) ? sizeof(cx<double>) : sizeof(cx<float>);
    p->n_phase = c->n_phase;
    p->n_atom = c->n_atom;
    for (int i = 0; i < 3; ++i) p->qC[i] = c->qC[i];
    const int nin[3] = {4, 3, 2};
    classify(p, grid, c);
    // pass 1 mixes the groups with the compile-time kernels or the run-time-radix
    // pass_strided_mix
    const int mix_role = p->nZ == 1 ? kRoleMix1 : (p->nZ == 2 ? kRoleMix2 : -1);
    p->grouped = (mix_role >= 0 && p->nS < p->nY && !class_form) ? 1 : 0;
    const int tab_entries = c->n_atom * c->n_phase * kAtomTab;
    int rc = QPD_OK;
    for (int d = 0; d < 3 && rc == QPD_OK; ++d) {
        const bool io32 = p->mixed != 0 && p->csz[d] == 16;
        rc = pass_candidates(p->fft[d], d, p->pd[d], p->dims, nin[d], p->csz[d], io32, true, tab_entries, true,
                             p->cand[d], p->grouped ? mix_role : -1, p->grouped ? p->nS : 0, p->grouped ? p->nY : 0);
        if (d == 1 && p->grouped && rc != QPD_OK) {
            // no source-group kernel (direct or split) for this length: keep the class form
            g_error.clear();
            p->grouped = 0;
            rc = pass_candidates(p->fft[d], d, p->pd[d], p->dims, nin[d], p->csz[d], io32, true, tab_entries, true,
                                 p->cand[d], -1);
            // pass 0 of the class form has no one-wave-per-row schedule
            std::vector<PassPlan>& c0 = p->cand[0];
            c0.erase(std::remove_if(c0.begin(), c0.end(),
                                    [](const PassPlan& pp) { return pp.L[0].fe && pp.L[0].fe->role == kRoleSourceW; }),
                     c0.end());
            p->sel[0] = 0;
            p->pass[0] = c0[0];
        }
        if (rc == QPD_OK) {
            // default: a kernel of the length itself, else the most even split, else generic
            p->sel[d] = 0;
            p->pass[d] = p->cand[d][0];
        }
    }
    // a split among the intensity pass's schedules of a plan computing in FP64 on FP32 spectra
    // needs the workspace of its FP64 intermediate
    if (rc == QPD_OK && p->mixed && p->csz[2] == 16)
        for (const PassPlan& c : p->cand[2]) p->split64 = p->split64 || c.nlaunch == 2;
    // the tables of each element type some pass computes in
    bool need32 = false, need64 = false;
    for (int d = 0; d < 3; ++d) (p->csz[d] == 16 ? need64 : need32) = true;
    if (rc == QPD_OK && need32) rc = build_tables<float>(p, grid, c, p->tab[0]);
    if (rc == QPD_OK && need64) rc = build_tables<double>(p, grid, c, p->tab[1]);
    if (rc != QPD_OK) {
        for (PlanTables& t : p->tab)
            if (t.dev) (void)hipFree(t.dev);
        delete p;
        return rc;
    }
    *plan = p;
    return QPD_OK;
}

int qpd_plan_info(const qpd_plan* p, int pass_dims[3], int* compute_bits, int* storage_bits, double* dlog10_estimate) {
    if (!p) return fail(QPD_ERR_ARG, "null plan");
    if (pass_dims)
        for (int d = 0; d < 3; ++d) pass_dims[d] = p->pd[d];
    if (compute_bits) *compute_bits = p->csz[0] == 16 ? 64 : 32;
    if (storage_bits) *storage_bits = p->ssz == 16 ? 64 : 32;
    if (dlog10_estimate) *dlog10_estimate = p->est_dlog10;
    return QPD_OK;
}

int qpd_plan_pass_bits(const qpd_plan* p, int bits[3]) {
    if (!p || !bits) return fail(QPD_ERR_ARG, "null argument");
    for (int d = 0; d < 3; ++d) bits[d] = p->csz[d] == 16 ? 64 : 32;
    return QPD_OK;
}

int qpd_plan_destroy(qpd_plan* plan) {
    if (!plan) return QPD_OK;
    if (plan->tab[0].dev || plan->tab[1].dev) (void)hipDeviceSynchronize();
    for (PlanTables& t : plan->tab)
        if (t.dev) (void)hipFree(t.dev);
    delete plan;
    return QPD_OK;
}

size_t qpd_workspace_bytes(const qpd_plan* p) {
    if (!p) return 0;
    return p->w1_bytes() + p->w2_bytes() + p->wide_bytes() + p->split64_bytes();
}


int qpd_ingest_fields(const qpd_ingest* in, int precision, void* P_out, double* U_out, signed char* phase_out,
                      void* stream) {
    g_error.clear();
    if (!in || !P_out || !U_out || !phase_out) return fail(QPD_ERR_ARG, "null argument");
    if (precision != 32 && precision != 64) return fail(QPD_ERR_ARG, "precision must be 32 or 64");
    if (in->P_bits != 32 && in->P_bits != 64) return fail(QPD_ERR_ARG, "P_bits must be 32 or 64");
    const bool has_u = in->U[0] != nullptr;
    if (has_u && (in->U_bits != 32 && in->U_bits != 64)) return fail(QPD_ERR_ARG, "U_bits must be 32 or 64");
    if (has_u && (!in->U[1] || !in->U[2])) return fail(QPD_ERR_ARG, "U needs all three components or none");
    if (!in->P[0] || !in->P[1] || !in->P[2] || !in->material || !in->dc || !in->phase_of)
        return fail(QPD_ERR_ARG, "null field pointer");
    for (int d = 0; d < 3; ++d)
        if (in->dims[d] < 1) return fail(QPD_ERR_ARG, "dims must be >= 1");
    if (in->z0 < 0 || in->z0 + in->dims[2] > in->nz_full) return fail(QPD_ERR_ARG, "the crop must lie inside the stack");
    if (in->n_material < 1) return fail(QPD_ERR_ARG, "n_material must be >= 1");
    if (!(in->P_div != 0.0) || (has_u && !(in->U_div != 0.0))) return fail(QPD_ERR_ARG, "zero unit divisor");
    const int64_t nc = in->dims[0] * in->dims[1], n = nc * in->dims[2];
    if (n >= (int64_t)1 << 31) return fail(QPD_ERR_UNSUPPORTED, "grids of 2^31 cells or more are not supported");
    IngestArgs a;
    a.nc = (uint32_t)nc;
    a.n = (uint32_t)n;
    a.nz_full = in->nz_full;
    a.z0 = in->z0;
    a.nz = in->dims[2];
    a.P0 = in->P[0];
    a.P1 = in->P[1];
    a.P2 = in->P[2];
    a.U0 = has_u ? in->U[0] : nullptr;
    a.U1 = has_u ? in->U[1] : nullptr;
    a.U2 = has_u ? in->U[2] : nullptr;
    a.P_div = in->P_div;
    a.U_div = has_u ? in->U_div : 1.0;
    a.dz = in->dz;
    a.mat = in->material;
    a.n_material = in->n_material;
    a.dc = in->dc;
    a.phase_of = in->phase_of;
    const int ub = has_u ? in->U_bits : 64;
    hipStream_t st = static_cast<hipStream_t>(stream);
    QPD_HIP(precision == 32 ? ingest_dispatch<float>(a, in->P_bits, ub, P_out, U_out, phase_out, st)
                            : ingest_dispatch<double>(a, in->P_bits, ub, P_out, U_out, phase_out, st));
    return QPD_OK;
}

int qpd_project(const int64_t dims[3], const double* A, double* S0, double* S1, double* S2, void* stream) {
    g_error.clear();
    if (!dims || !A) return fail(QPD_ERR_ARG, "null argument");
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 1) return fail(QPD_ERR_ARG, "dims must be >= 1");
    if (dims[0] * dims[1] * dims[2] >= (int64_t)1 << 31)
        return fail(QPD_ERR_UNSUPPORTED, "arrays of 2^31 values or more are not supported");
    const uint32_t d0 = (uint32_t)dims[0], d1 = (uint32_t)dims[1], d2 = (uint32_t)dims[2];
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (S0) {
        const uint32_t rows = d1 * d2;
        hipLaunchKernelGGL(project_inner, dim3((rows + 3u) / 4u), dim3(256), 0, st, A, S0, d0, rows);
        QPD_HIP(hipGetLastError());
    }
    if (S1) {
        hipLaunchKernelGGL(project_middle, dim3((d0 * d2 + 255u) / 256u), dim3(256), 0, st, A, S1, d0, d1, d2);
        QPD_HIP(hipGetLastError());
    }
    if (S2) {
        hipLaunchKernelGGL(project_outer, dim3((d0 * d1 + 255u) / 256u), dim3(256), 0, st, A, S2, d0, d1, d2);
        QPD_HIP(hipGetLastError());
    }
    return QPD_OK;
}

int qpd_reverse_axes(const int64_t dims[3], int n_comp, int elem_bytes, const void* src, void* dst, void* stream) {
    g_error.clear();
    if (!dims || !src || !dst) return fail(QPD_ERR_ARG, "null argument");
    if (n_comp < 1) return fail(QPD_ERR_ARG, "n_comp must be >= 1");
    for (int d = 0; d < 3; ++d)
        if (dims[d] < 1) return fail(QPD_ERR_ARG, "dims must be >= 1");
    const int64_t n = dims[0] * dims[1] * dims[2];
    if (n >= (int64_t)1 << 31) return fail(QPD_ERR_UNSUPPORTED, "grids of 2^31 cells or more are not supported");
    const char* s = static_cast<const char*>(src);
    const char* d = static_cast<const char*>(dst);
    const int64_t bytes = n * n_comp * elem_bytes;
    if (s < d + bytes && d < s + bytes) return fail(QPD_ERR_ARG, "src and dst overlap");
    const uint32_t n0 = (uint32_t)dims[0], n1 = (uint32_t)dims[1], n2 = (uint32_t)dims[2];
    const uint32_t t0 = (n0 + kTrTile - 1) / kTrTile, t2 = (n2 + kTrTile - 1) / kTrTile;
    const uint64_t blocks = (uint64_t)t0 * t2 * n1 * (uint64_t)n_comp;
    if (blocks > 0x7fffffffULL) return fail(QPD_ERR_UNSUPPORTED, "too many tiles for one launch");
    hipStream_t st = static_cast<hipStream_t>(stream);
    const dim3 grid((unsigned)blocks), block(kTrTile, kTrRows);
    if (elem_bytes == 1)
        hipLaunchKernelGGL(reverse_axes_kernel<uint8_t>, grid, block, 0, st, static_cast<const uint8_t*>(src),
                           static_cast<uint8_t*>(dst), n0, n1, n2, t0, t2);
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(reverse_axes_kernel<uint32_t>, grid, block, 0, st, static_cast<const uint32_t*>(src),
                           static_cast<uint32_t*>(dst), n0, n1, n2, t0, t2);
    else if (elem_bytes == 8)
        hipLaunchKernelGGL(reverse_axes_kernel<uint64_t>, grid, block, 0, st, static_cast<const uint64_t*>(src),
                           static_cast<uint64_t*>(dst), n0, n1, n2, t0, t2);
    else
        return fail(QPD_ERR_ARG, "elem_bytes must be 1, 4 or 8");
    QPD_HIP(hipGetLastError());
    return QPD_OK;
}

int qpd_fast_lengths(int role, int64_t* lengths, int cap) {
    if (role < 0 || role >= kNumRoles) return fail(QPD_ERR_ARG, "role must be 0..12");
    std::vector<int64_t> ls;
    for (const FastEntry& e : fast_registry())
        if (e.role == role && e.real_bytes == 4 && std::find(ls.begin(), ls.end(), (int64_t)e.N) == ls.end())
            ls.push_back(e.N);
    std::sort(ls.begin(), ls.end());
    for (int i = 0; i < cap && i < (int)ls.size() && lengths; ++i) lengths[i] = ls[i];
    return (int)ls.size();
}

int qpd_plan_schedules(const qpd_plan* p, int pass) {
    if (!p || pass < 0 || pass > 2) return fail(QPD_ERR_ARG, "null plan or pass not 0, 1, 2");
    return (int)p->cand[pass].size();
}

int qpd_plan_schedule_name(const qpd_plan* p, int pass, int index, char* buf, int cap) {
    if (!p || pass < 0 || pass > 2) return fail(QPD_ERR_ARG, "null plan or pass not 0, 1, 2");
    if (index < 0 || index >= (int)p->cand[pass].size()) return fail(QPD_ERR_ARG, "schedule index out of range");
    const std::string name = describe(p->cand[pass][index]);
    if (buf && cap > 0) {
        const size_t n = std::min((size_t)cap - 1, name.size());
        std::memcpy(buf, name.data(), n);
        buf[n] = 0;
    }
    return (int)name.size();
}

int qpd_plan_set_schedule(qpd_plan* p, int pass, int index) {
    if (!p || pass < 0 || pass > 2) return fail(QPD_ERR_ARG, "null plan or pass not 0, 1, 2");
    if (index < 0 || index >= (int)p->cand[pass].size()) return fail(QPD_ERR_ARG, "schedule index out of range");
    p->sel[pass] = index;
    p->pass[pass] = p->cand[pass][index];
    return QPD_OK;
}

int qpd_plan_get_schedule(const qpd_plan* p, int pass) {
    if (!p || pass < 0 || pass > 2) return fail(QPD_ERR_ARG, "null plan or pass not 0, 1, 2");
    return p->sel[pass];
}

int qpd_launches_per_volume(const qpd_plan* p) {
    return p ? p->pass[0].nlaunch + p->pass[1].nlaunch + p->pass[2].nlaunch : 0;
}

int qpd_intensity_pass(qpd_plan* p, const qpd_fields* f, double* I, void* work, int pass, void* stream) {
    if (!p || !f || !I || !work) return fail(QPD_ERR_ARG, "null argument");
    if (!f->P || !f->U || !f->phase) return fail(QPD_ERR_ARG, "P, U and phase are required");
    if (pass < 0 || pass > 2) return fail(QPD_ERR_ARG, "pass must be 0, 1 or 2");
    hipStream_t st = static_cast<hipStream_t>(stream);
    return run_intensity(p, f, I, work, st, pass);
}

int qpd_intensity_noalloc(qpd_plan* p, const qpd_fields* f, double* I, void* stream) {
    if (!p || !f || !I) return fail(QPD_ERR_ARG, "null argument");
    if (!f->P || !f->U || !f->phase) return fail(QPD_ERR_ARG, "P, U and phase are required");
    const size_t need = qpd_workspace_bytes(p);
    if (need > QPD_NOALLOC_SCRATCH_BYTES)
        return fail(QPD_ERR_UNSUPPORTED, "the plan needs " + std::to_string(need) +
                                             " bytes of workspace; qpd_intensity_noalloc holds at most " +
                                             std::to_string(QPD_NOALLOC_SCRATCH_BYTES) +
                                             " (use qpd_intensity with a workspace)");
    void* work = nullptr;
    const int rc = scratch_of_current_device(&work);
    if (rc != QPD_OK) return rc;
    return run_intensity(p, f, I, work, static_cast<hipStream_t>(stream));
}

double qpd_pass_bytes_per_cell(const qpd_plan* p, const qpd_fields* f, int pass) {
    if (!p || pass < 0 || pass > 2) return 0.0;
    // a wide plan reads each FP32 input, writes it widened and reads that in pass 0
    const double real = p->precision == 32 ? (p->wide ? 4.0 + 8.0 + 8.0 : 4.0) : 8.0;
    const double c0 = (double)p->ssz, c1 = (double)p->ssz;
    if (pass == 0) {
        double in = 3 * real + 3 * 8.0 + 1.0;  // P, U, phase
        if (f && f->region) in += real;
        if (f && f->fraction) in += p->n_phase * real;
        return in + (double)p->n1fields() * c0;
    }
    if (pass == 1) return (double)p->n1fields() * c0 + 2.0 * p->nZ * c1;
    return 2.0 * p->nZ * c1 + 8.0;
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
    return run_intensity(p, f, I, work, st);
}

}  // extern "C"
