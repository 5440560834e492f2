// Launch shapes and the registry of compile-time-length kernels (host side).
//
// A pass is one or two kernel launches.  Each launch is either a generic kernel
// (run-time radices, qpd_kernels.hpp) or an entry of the fast registry: a kernel
// of qpd_fast.hpp instantiated for one transform length, one role and one tile
// width G.  The registry is filled from several translation units (one per group
// of lengths) so that the library builds in parallel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "qpd_kernels.hpp"

namespace qpd {

// Roles of a fast kernel.
enum FastRole {
    kRoleSource = 0,   // contiguous pass with the source: 4 inputs per atom -> 3 outputs per class
    kRoleContig = 1,   // contiguous plain transform (one input per field)
    kRoleStrided = 2,  // strided pass, 3 inputs -> 2 outputs per class
    kRoleIntensity = 3,  // strided pass, 2 inputs -> fftshifted intensity
    kRoleStrided1 = 4,   // strided plain transform
    kRoleSplit = 5,      // strided length-N2 sub-transforms of a split length, with the twiddle, in place
    kRoleStridedB = 6,   // kRoleStrided as the second step of a split length
    kRoleIntensityB = 7, // kRoleIntensity as the second step of a split length
    kRoleMix1 = 8,       // strided pass 1 over source groups, one z-class
    kRoleMix2 = 9,       // strided pass 1 over source groups, two z-classes
    kRoleMix1B = 10,     // kRoleMix1 as the second step of a split length
    kRoleMix2B = 11,     // kRoleMix2 as the second step of a split length
    kRoleSourceW = 12,   // contiguous pass with the source over source groups, one wave per row
    kNumRoles = 13,
};

struct FastEntry;

// Launch shape of one kernel of a pass.
struct Launch {
    const FastEntry* fe = nullptr;  // fast kernel, or nullptr for a generic kernel
    int contiguous = 0;             // generic: 1 pass_contig, 0 pass_strided
    int nl = 0;                     // rows (contig) or TX (strided) per block
    int nf = 0;                     // inputs per LDS fill (generic)
    int nthr = 0;
    int ept = 0;                    // elements per thread, 4 / 8 (generic)
    int rset = 0;                   // generic FP64: the radix set of the axis (kRadixAll, kRadixPow2, kRadixSmall)
    int io32 = 0;                   // generic FP64 kernel on FP32 arrays (FP32 plan computing in FP64)
    int mix = 0;                    // generic strided pass 1 over source groups (pass_strided_mix)
    size_t lds = 0;
    int64_t blocks = 0;
};

// One pass of a plan: one launch, or two for a split strided transform of length
// N = N1 N2 (the first launch runs the N2-point sub-transforms with the twiddle in
// place, the second the N1-point transforms with the pass's accumulation).
struct PassPlan {
    int nlaunch = 1;
    int N1 = 0, N2 = 0;  // split lengths, 0 when the pass is not split
    Launch L[2];
};

using FastLaunchFn = hipError_t (*)(const Launch&, const PassArgs<float>&, hipStream_t);
using FastLaunchFn64 = hipError_t (*)(const Launch&, const PassArgs<double>&, hipStream_t);

struct FastEntry {
    int N;
    int role;
    int G;        // rows (contiguous) or columns (strided) per block
    int nthr;
    int tpl;      // threads per line
    size_t lds;   // dynamic LDS bytes, without the atom table of the source role
    const char* radices;
    FastLaunchFn launch;
    const void* kernel;  // the kernel function (for its attributes)
    bool no_prefetch;    // strided roles: the variant without the register prefetch
    int real_bytes = 4;  // 4: an FP32 kernel (launch), 8: an FP64 one (launch64)
    FastLaunchFn64 launch64 = nullptr;
    bool io32 = false;   // FP64 kernel on FP32 arrays (P, weights, spectra): FP32 plans computing in FP64
    bool one_buffer = false;  // strided FP64 kernel with one LDS stage buffer
    int nsrc = 0;             // pass 1 over source groups: the group count it requires (0: any)
    int split = 1;            // one-wave-per-row source pass: lanes per last-stage butterfly
    bool lds_stage = false;   // inputs staged in LDS by sixteen-byte loads: even n0 and strides
    bool tab_lds = false;     // pass over source groups with its tables in LDS (a.ncoef (N + a.nsrc G) values more)
};

// Raise a kernel's dynamic LDS limit once per (kernel, device) and size.
hipError_t allow_lds(const void* kernel, int lds);

template <typename A>
hipError_t launch_one(void (*kernel)(A), const Launch& L, hipStream_t st, const A& a) {
    if (L.lds > 65536) {
        hipError_t e = allow_lds((const void*)kernel, (int)L.lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)L.blocks), dim3((unsigned)L.nthr), L.lds, st, a);
    return hipGetLastError();
}

// Every registered fast kernel (filled once, on first use).
const std::vector<FastEntry>& fast_registry();

// qpd_fast_sizes.hip is compiled once per part (QPD_FAST_PART = 0 .. kFastParts - 1);
// part p instantiates every kFastParts-th length of its list and registers it here.
constexpr int kFastParts = 8;
void register_fast_part0(std::vector<FastEntry>& v);
void register_fast_part1(std::vector<FastEntry>& v);
void register_fast_part2(std::vector<FastEntry>& v);
void register_fast_part3(std::vector<FastEntry>& v);
void register_fast_part4(std::vector<FastEntry>& v);
void register_fast_part5(std::vector<FastEntry>& v);
void register_fast_part6(std::vector<FastEntry>& v);
void register_fast_part7(std::vector<FastEntry>& v);

}  // namespace qpd
