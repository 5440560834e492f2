// Instantiation helpers of the fast-kernel registry (included by qpd_fast_sizes.hip only).
#pragma once

#include "qpd_fast.hpp"
#include "qpd_launch.hpp"

namespace qpd {

// Default tile width G of a role for a length with TPL threads per line: the source
// pass runs 4 G TPL threads (about 256), the plain contiguous pass about 256; strided
// passes take 16 columns (128-byte rows of FP32 complex) while the block stays within
// 512 threads, else 8, else 4 (0: no strided kernel for this length).  Blocks of at
// most 512 threads leave each wave 256 VGPRs.
constexpr int fast_tile(int role, int tpl) {
    return role == kRoleSource   ? (tpl >= 64 ? 1 : 64 / tpl)
           : role == kRoleContig ? (tpl >= 256 ? 1 : 256 / tpl)
                                 : (16 * tpl <= 512 ? 16 : (8 * tpl <= 512 ? 8 : (4 * tpl <= 512 ? 4 : 0)));
}

template <class SZ, int ROLE, int G, bool PF = true>
hipError_t fast_launch(const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    if constexpr (ROLE == kRoleSource)
        return launch_one(fast_source<float, SZ, G>, L, st, a);
    else if constexpr (ROLE == kRoleSourceW)
        return launch_one(fast_source_wave<float, SZ, G / kSrcWaves>, L, st, a);
    else if constexpr (ROLE == kRoleContig)
        return launch_one(fast_contig<float, SZ, G, 1, false>, L, st, a);
    else if constexpr (ROLE == kRoleStrided)
        return launch_one(fast_strided<float, SZ, G, 3, false, false>, L, st, a);
    else if constexpr (ROLE == kRoleIntensity)
        return launch_one(fast_strided<float, SZ, G, 2, true, false, PF>, L, st, a);
    else if constexpr (ROLE == kRoleStrided1)
        return launch_one(fast_strided<float, SZ, G, 1, false, false>, L, st, a);
    else if constexpr (ROLE == kRoleSplit)
        return launch_one(fast_split<float, SZ, G>, L, st, a);
    else if constexpr (ROLE == kRoleStridedB)
        return launch_one(fast_strided<float, SZ, G, 3, false, true>, L, st, a);
    else if constexpr (ROLE == kRoleIntensityB)
        return launch_one(fast_strided<float, SZ, G, 2, true, true>, L, st, a);
    else if constexpr (ROLE == kRoleMix1)
        return launch_one(fast_strided_mix<float, SZ, G, 1, false>, L, st, a);
    else if constexpr (ROLE == kRoleMix2)
        return launch_one(fast_strided_mix<float, SZ, G, 2, false>, L, st, a);
    else if constexpr (ROLE == kRoleMix1B)
        return launch_one(fast_strided_mix<float, SZ, G, 1, true>, L, st, a);
    else
        return launch_one(fast_strided_mix<float, SZ, G, 2, true>, L, st, a);
}

// The first step of a split strided pass with the next field staged in LDS (fast_split
// LDSIN): FP32, and FP64 on FP32 arrays; listed before the register-prefetch kernels.
template <class SZ, int G>
hipError_t split_lds_launch(const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    return launch_one(fast_split<float, SZ, G, false, true>, L, st, a);
}
template <class SZ, int G>
hipError_t split_lds_launch64(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    return launch_one(fast_split<double, SZ, G, true, true>, L, st, a);
}
template <class SZ, int G, bool F64>
void add_split_lds(std::vector<FastEntry>& v, const char* radices) {
    if constexpr (G >= 2 && G <= 16 && G * SZ::TPL >= 64 && G * SZ::TPL <= (F64 ? 256 : 512)) {
        constexpr size_t cs = F64 ? 16 : 8;
        constexpr size_t dyn = ((size_t)SZ::strided_tile(G) + SZ::N) * cs;
        constexpr int RPI = 64 / (G / 2);
        constexpr size_t stat = (size_t)((SZ::N + RPI - 1) / RPI) * RPI * G * 8 + (size_t)SZ::N * cs;
        if constexpr (dyn + stat <= 160 * 1024) {
            FastEntry e{SZ::N, kRoleSplit, G, G * SZ::TPL, SZ::TPL, dyn, radices,
                        F64 ? nullptr : &split_lds_launch<SZ, G>,
                        F64 ? (const void*)fast_split<double, SZ, G, true, true> : (const void*)fast_split<float, SZ, G, false, true>,
                        false, F64 ? 8 : 4, F64 ? &split_lds_launch64<SZ, G> : nullptr, F64, true};
            e.lds_stage = true;
            v.push_back(e);
        }
    }
}

// The pass over source groups (unsplit or the second step of a split) with its tables in LDS
// (fast_strided_mix TABL), FP32; listed before the kernel that reads them from global memory.
template <class SZ, int ROLE, int G>
hipError_t mix_tabl_launch(const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    constexpr int NZ = (ROLE == kRoleMix1 || ROLE == kRoleMix1B) ? 1 : 2;
    constexpr bool SPLIT = ROLE == kRoleMix1B || ROLE == kRoleMix2B;
    return launch_one(fast_strided_mix<float, SZ, G, NZ, SPLIT, false, false, true>, L, st, a);
}
template <class SZ, int ROLE, int G>
void add_mix_tabl(std::vector<FastEntry>& v, const char* radices) {
    constexpr int NZ = (ROLE == kRoleMix1 || ROLE == kRoleMix1B) ? 1 : 2;
    constexpr bool SPLIT = ROLE == kRoleMix1B || ROLE == kRoleMix2B;
    constexpr size_t lds = (2 * (size_t)SZ::strided_tile(G) + SZ::N + (SZ::N + 1) / 2) * 8;
    if constexpr (G >= 1 && G * SZ::TPL <= 1024 && lds <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, lds, radices, &mix_tabl_launch<SZ, ROLE, G>,
                    (const void*)fast_strided_mix<float, SZ, G, NZ, SPLIT, false, false, true>, false};
        e.tab_lds = true;
        v.push_back(e);
    }
}

template <class SZ, int ROLE, int G, bool PF = true>
void add_fast(std::vector<FastEntry>& v, const char* radices) {
    if constexpr (G >= 1) {
        constexpr int F = ROLE == kRoleSource ? 4 : 1;
        constexpr int nthr = F * G * SZ::TPL;
        // two stage buffers and the twiddle table
        constexpr size_t tile = ROLE <= kRoleContig ? (size_t)F * G * SZ::PITCH : (size_t)SZ::strided_tile(G);
        constexpr size_t lds = (2 * tile + SZ::N) * 8;
        if constexpr (nthr <= 1024 && lds <= 160 * 1024)
            v.push_back({SZ::N, ROLE, G, nthr, SZ::TPL, lds, radices, &fast_launch<SZ, ROLE, G, PF>, nullptr, !PF});
    }
}

// FP64 kernels of the strided roles and the block-per-row source pass: the same code over
// complex FP64.  A thread holds the same points as in FP32 at twice the registers, so the
// strided tiles take half the columns (rows of the same bytes), and every block has at most
// 256 threads: a wave may then use 512 registers (AGPRs included) instead of spilling.
// IO32: the same kernels computing in FP64 on FP32 arrays (P, weights, the spectra between
// passes), for FP32 plans whose transforms run in FP64.
template <class SZ, int ROLE, int G, bool PF = true, bool IO32 = false, bool ONEBUF = false>
hipError_t fast_launch64(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    if constexpr (ROLE == kRoleSource)
        return launch_one(fast_source<double, SZ, G, IO32>, L, st, a);
    else if constexpr (ROLE == kRoleSourceW)
        return launch_one(fast_source_wave<double, SZ, G / kSrcWaves, IO32>, L, st, a);
    else if constexpr (ROLE == kRoleStrided)
        return launch_one(fast_strided<double, SZ, G, 3, false, false, true, IO32, ONEBUF>, L, st, a);
    else if constexpr (ROLE == kRoleIntensity)
        return launch_one(fast_strided<double, SZ, G, 2, true, false, PF, IO32, ONEBUF>, L, st, a);
    else if constexpr (ROLE == kRoleSplit)
        return launch_one(fast_split<double, SZ, G, IO32>, L, st, a);
    else if constexpr (ROLE == kRoleIntensityB)
        return launch_one(fast_strided<double, SZ, G, 2, true, true, PF, IO32>, L, st, a);
    else if constexpr (ROLE == kRoleMix1)
        return launch_one(fast_strided_mix<double, SZ, G, 1, false, IO32, ONEBUF>, L, st, a);
    else
        return launch_one(fast_strided_mix<double, SZ, G, 2, false, IO32, ONEBUF>, L, st, a);
}

// The FP64 intensity pass (IO32: on FP32 spectra) or the second step of its split (on the
// FP64 intermediate) with the next input staged in LDS (fast_strided LDSIN, one LDS buffer).
template <class SZ, int ROLE, int G, bool IO32>
hipError_t strided_lds_launch64(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    return launch_one(fast_strided<double, SZ, G, 2, true, ROLE == kRoleIntensityB, false, IO32, true, true>, L, st, a);
}
template <class SZ, int ROLE, int G, bool IO32>
void add_strided_lds(std::vector<FastEntry>& v, const char* radices) {
    constexpr size_t ce = IO32 ? 8 : 16;                       // bytes of a staged element
    constexpr int EPL = 16 / (int)ce, LPR = G / EPL > 0 ? G / EPL : 1, RPI = 64 / LPR;
    constexpr size_t dyn = ((size_t)SZ::strided_tile(G) + SZ::N) * 16;
    constexpr size_t stat = (size_t)((SZ::N + RPI - 1) / RPI) * RPI * G * ce;
    if constexpr (G >= 1 && G % EPL == 0 && G * SZ::TPL >= 64 && G * SZ::TPL <= 256 && dyn + stat <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, dyn, radices, nullptr,
                    (const void*)fast_strided<double, SZ, G, 2, true, ROLE == kRoleIntensityB, false, IO32, true, true>,
                    true, 8, &strided_lds_launch64<SZ, ROLE, G, IO32>, IO32, true};
        e.lds_stage = true;
        v.push_back(e);
    }
}

// ONEBUF (strided roles): one LDS stage buffer instead of two
template <class SZ, int ROLE, int G, bool PF = true, bool IO32 = false, bool ONEBUF = false>
void add_fast64(std::vector<FastEntry>& v, const char* radices) {
    constexpr bool src = ROLE == kRoleSource;
    if constexpr (G >= (src ? 1 : 4) && (src ? 4 : 1) * G * SZ::TPL <= 256 && !(src && ONEBUF)) {
        constexpr size_t tile = src ? (size_t)4 * G * SZ::PITCH : (size_t)SZ::strided_tile(G);
        constexpr size_t lds = ((ONEBUF ? 1 : 2) * tile + SZ::N) * sizeof(cx<double>);
        if constexpr (lds <= 160 * 1024)
            v.push_back({SZ::N, ROLE, G, (src ? 4 : 1) * G * SZ::TPL, SZ::TPL, lds, radices, nullptr, nullptr, !PF,
                         8, &fast_launch64<SZ, ROLE, G, PF, IO32, ONEBUF>, IO32, ONEBUF});
    }
}

// FP32 plans computing in FP64 (IO32): the intensity pass lists its tiles with one LDS
// stage buffer first (the two-buffer tile of a 512-point pass, 82 KB, leaves one block of
// four waves on a CU, one buffer, 45 KB, two: 0.72 -> 0.46 ms at 512x512x104); the plain
// strided pass lists it after the two-buffer tile.  The pass over source groups holds
// about 390 registers a thread, one wave per SIMD whatever its LDS, and keeps two buffers.
template <class SZ, int ROLE, bool IO32>
void add_fast64_tiles(std::vector<FastEntry>& v, const char* radices) {
    constexpr int G = fast_tile(ROLE, SZ::TPL) / (ROLE == kRoleSource ? 1 : 2);
    if constexpr (ROLE == kRoleIntensity) {
        if constexpr (IO32) {
            add_fast64<SZ, ROLE, G, false, IO32, true>(v, radices);
            add_strided_lds<SZ, ROLE, G, true>(v, radices);
            add_fast64<SZ, ROLE, G, true, IO32, true>(v, radices);
        }
        add_fast64<SZ, ROLE, G, true, IO32>(v, radices);
        add_fast64<SZ, ROLE, G, false, IO32>(v, radices);
        add_fast64<SZ, ROLE, G / 2, false, IO32>(v, radices);
    } else {
        add_fast64<SZ, ROLE, G, true, IO32>(v, radices);
        if constexpr (IO32 && ROLE == kRoleStrided) add_fast64<SZ, ROLE, G, true, IO32, true>(v, radices);
        if constexpr (ROLE != kRoleSource) add_fast64<SZ, ROLE, G / 2, true, IO32>(v, radices);
    }
}

// FP64 plans, and FP32 plans computing in FP64 (FP32 arrays, IO32)
template <class SZ, int ROLE>
void add_fast64_variants(std::vector<FastEntry>& v, const char* radices) {
    add_fast64_tiles<SZ, ROLE, false>(v, radices);
    add_fast64_tiles<SZ, ROLE, true>(v, radices);
}

// Pass 1 over exactly NS source groups with the group spectra held (fast_mix_held), listed
// before the general pass over source groups: FP32 at the strided tile, FP64 (and FP64 on
// FP32 arrays) at half of it.
template <class SZ, int NZ, int NS, int G, bool PF = true>
hipError_t held_launch(const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    return launch_one(fast_mix_held<float, SZ, G, NZ, NS, false, false, PF>, L, st, a);
}
template <class SZ, int NZ, int NS, int G, bool IO32, bool ONEBUF = false, bool PF = true>
hipError_t held_launch64(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    return launch_one(fast_mix_held<double, SZ, G, NZ, NS, IO32, ONEBUF, PF>, L, st, a);
}

template <class SZ, int NZ, int NS, int G, int WPE>
hipError_t held_lds_launch(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    return launch_one(fast_mix_held_lds<SZ, G, NZ, NS, WPE>, L, st, a);
}
template <class SZ, int NZ, int NS, int G>
hipError_t held_lds_launch64w(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    return launch_one(fast_mix_held_lds<SZ, G, NZ, NS, 2, double, false>, L, st, a);
}
template <class SZ, int NZ, int NS, int G>
hipError_t held_lds_launch32(const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    return launch_one(fast_mix_held_lds<SZ, G, NZ, NS, 1, float>, L, st, a);
}

template <class SZ, int ROLE, int NS>
void add_held(std::vector<FastEntry>& v, const char* radices) {
    constexpr int NZ = ROLE == kRoleMix1 ? 1 : 2;
    constexpr int G = fast_tile(ROLE, SZ::TPL);
    constexpr size_t tile = (size_t)SZ::strided_tile(G), tile64 = (size_t)SZ::strided_tile(G / 2);
    constexpr size_t tabs = SZ::N + (SZ::N + 1) / 2;  // twiddles and wavenumbers
    // FP32, first with the next input staged in LDS (one LDS stage buffer; two groups, an even
    // number of columns): 0.546 ms against 0.582 without a prefetch and 0.601 with a register prefetch
    // for the 104-point pass of a 512 x 512 x 104 volume
    if constexpr (NS == 2 && G == 16 && G * SZ::TPL >= 64 && (tile + tabs) * 8 <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, (tile + tabs) * 8, radices, &held_lds_launch32<SZ, NZ, NS, G>,
                    (const void*)fast_mix_held_lds<SZ, G, NZ, NS, 1, float>, false};
        e.nsrc = NS;
        e.one_buffer = true;
        e.lds_stage = true;
        v.push_back(e);
    }
    if constexpr (G >= 1 && (2 * tile + tabs) * 8 <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, (2 * tile + tabs) * 8, radices,
                    &held_launch<SZ, NZ, NS, G>, (const void*)fast_mix_held<float, SZ, G, NZ, NS>, false};
        e.nsrc = NS;
        v.push_back(e);
        // without the register prefetch: 106 registers, four waves per SIMD instead of three
        // (0.588 against 0.603 ms for the 104-point pass of a 512 x 512 x 104 volume)
        FastEntry f{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, (2 * tile + tabs) * 8, radices,
                    &held_launch<SZ, NZ, NS, G, false>, (const void*)fast_mix_held<float, SZ, G, NZ, NS, false, false, false>,
                    true};
        f.nsrc = NS;
        v.push_back(f);
    }
    // FP64 on FP32 arrays at the FP32 tile (rows of 128 bytes read and written, blocks up to
    // 256 threads), first with the next input staged in LDS (two groups, whose held values fit
    // two waves per SIMD; an even number of columns; the 104-point pass of a 512 x 512 x 104 volume:
    // 0.708 ms against 0.773 with one LDS buffer and a register prefetch, 0.785 with two
    // buffers, 0.827 at eight columns)
    if constexpr (NS == 2 && G == 16 && G * SZ::TPL >= 64 && G * SZ::TPL <= 256 && (tile + tabs) * 16 <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, (tile + tabs) * 16, radices, nullptr,
                    (const void*)fast_mix_held_lds<SZ, G, NZ, NS, 2>, false, 8, &held_lds_launch<SZ, NZ, NS, G, 2>, true,
                    true};
        e.nsrc = NS;
        e.lds_stage = true;
        v.push_back(e);
    }
    // one LDS stage buffer (a barrier before each stage write too): half the LDS, for lengths
    // whose blocks LDS rather than registers limits per CU
    if constexpr (G >= 1 && G * SZ::TPL <= 256 && (tile + tabs) * 16 <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, (tile + tabs) * 16, radices, nullptr,
                    (const void*)fast_mix_held<double, SZ, G, NZ, NS, true, true>, false, 8,
                    &held_launch64<SZ, NZ, NS, G, true, true>, true, true};
        e.nsrc = NS;
        v.push_back(e);
    }
    if constexpr (G >= 1 && G * SZ::TPL <= 256 && (2 * tile + tabs) * 16 <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G, G * SZ::TPL, SZ::TPL, (2 * tile + tabs) * 16, radices, nullptr,
                    (const void*)fast_mix_held<double, SZ, G, NZ, NS, true>, false, 8,
                    &held_launch64<SZ, NZ, NS, G, true>, true};
        e.nsrc = NS;
        v.push_back(e);
    }
    // FP64 plans, first with the next input staged in LDS (FP64 spectra: one value a lane, rows
    // of 128 bytes at half the FP32 tile; 192 registers against 244): 1.033 against 1.035 ms for
    // the 104-point pass of a 512 x 512 x 104 volume, 0.147 against 0.159 for the 64-point pass
    // of a 256 x 256 x 64 volume
    if constexpr (NS == 2 && G == 16 && (G / 2) * SZ::TPL >= 64 && (G / 2) * SZ::TPL <= 256 &&
                  (tile64 + tabs) * 16 <= 160 * 1024) {
        FastEntry e{SZ::N, ROLE, G / 2, (G / 2) * SZ::TPL, SZ::TPL, (tile64 + tabs) * 16, radices, nullptr,
                    (const void*)fast_mix_held_lds<SZ, G / 2, NZ, NS, 2, double, false>, false, 8,
                    &held_lds_launch64w<SZ, NZ, NS, G / 2>, false, true};
        e.nsrc = NS;
        e.lds_stage = true;
        v.push_back(e);
    }
    // FP64 (and FP64 on FP32 arrays) at half the FP32 tile
    if constexpr (G / 2 >= 4 && (G / 2) * SZ::TPL <= 256 && (2 * tile64 + tabs) * 16 <= 160 * 1024) {
        for (const bool io32 : {true, false}) {
            if (!io32 && NS > 2) continue;
            FastEntry e{SZ::N, ROLE, G / 2, (G / 2) * SZ::TPL, SZ::TPL, (2 * tile64 + tabs) * 16, radices, nullptr,
                        io32 ? (const void*)fast_mix_held<double, SZ, G / 2, NZ, NS, true>
                             : (const void*)fast_mix_held<double, SZ, G / 2, NZ, NS, false>,
                        false, 8, io32 ? &held_launch64<SZ, NZ, NS, G / 2, true> : &held_launch64<SZ, NZ, NS, G / 2, false>,
                        io32};
            e.nsrc = NS;
            v.push_back(e);
        }
    }
}

// The largest P <= TPL / NRL for which the last stage of a two-stage length splits over P
// lanes per butterfly (a prime last radix R with (R - 1) divisible by 2 P), else 1.
template <class SZ>
constexpr int split_lanes() {
    constexpr int R = SZ::RL;
    bool prime = R >= 5;
    for (int d = 2; d * d <= R; ++d) prime = prime && R % d != 0;
    if (SZ::NST != 2 || !prime) return 1;
    for (int p = SZ::TPL / SZ::NRL; p >= 2; --p)
        if ((R - 1) % (2 * p) == 0) return p;
    return 1;
}

template <class SZ, int ROWS, int P>
hipError_t srcw_launch(const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    return launch_one(fast_source_wave<float, SZ, ROWS, false, P>, L, st, a);
}
template <class SZ, int ROWS, bool IO32, int P>
hipError_t srcw_launch64(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    return launch_one(fast_source_wave<double, SZ, ROWS, IO32, P>, L, st, a);
}

// Threads per butterfly of the block-per-row source pass's first stage (a prime radix R0 with
// nr(0) butterflies a line and TPL threads a line): the largest P0 with P0 nr(0) <= TPL and
// (R0 - 1) divisible by 2 P0, else 1.
template <class SZ>
constexpr int first_split_lanes() {
    constexpr int R = SZ::rad(0);
    bool prime = R >= 5;
    for (int d = 2; d * d <= R; ++d) prime = prime && R % d != 0;
    if (!prime) return 1;
    for (int p = SZ::TPL / SZ::nr(0); p >= 2; --p)
        if ((R - 1) % (2 * p) == 0) return p;
    return 1;
}

template <class SZ, int G, int P0>
hipError_t src_split_launch(const Launch& L, const PassArgs<float>& a, hipStream_t st) {
    return launch_one(fast_source<float, SZ, G, false, P0>, L, st, a);
}
template <class SZ, int G, bool IO32, int P0>
hipError_t src_split_launch64(const Launch& L, const PassArgs<double>& a, hipStream_t st) {
    return launch_one(fast_source<double, SZ, G, IO32, P0>, L, st, a);
}

// The block-per-row source pass with its first stage split over P0 threads per butterfly
// (first_split_lanes): FP32 at its default rows a block and twice them, FP64 and FP64 on FP32
// arrays at the rows the 256-thread FP64 blocks allow; listed before the unsplit forms.
template <class SZ>
void add_source_split(std::vector<FastEntry>& v, const char* radices) {
    constexpr int P0 = first_split_lanes<SZ>();
    if constexpr (P0 > 1) {
        constexpr int G = fast_tile(kRoleSource, SZ::TPL);
        auto add32 = [&](auto g_) {
            constexpr int GG = decltype(g_)::value;
            constexpr size_t lds = (2 * (size_t)4 * GG * SZ::PITCH + SZ::N) * 8;
            if constexpr (4 * GG * SZ::TPL <= 512 && lds <= 160 * 1024) {
                FastEntry e{SZ::N, kRoleSource, GG, 4 * GG * SZ::TPL, SZ::TPL, lds, radices,
                            &src_split_launch<SZ, GG, P0>, (const void*)fast_source<float, SZ, GG, false, P0>, false};
                e.split = P0;
                v.push_back(e);
            }
        };
        add32(std::integral_constant<int, G>{});
        add32(std::integral_constant<int, 2 * G>{});
        constexpr size_t lds64 = (2 * (size_t)4 * G * SZ::PITCH + SZ::N) * 16;
        if constexpr (4 * G * SZ::TPL <= 256 && lds64 <= 160 * 1024) {
            for (const bool io32 : {true, false}) {
                FastEntry e{SZ::N, kRoleSource, G, 4 * G * SZ::TPL, SZ::TPL, lds64, radices, nullptr,
                            io32 ? (const void*)fast_source<double, SZ, G, true, P0>
                                 : (const void*)fast_source<double, SZ, G, false, P0>,
                            false, 8, io32 ? &src_split_launch64<SZ, G, true, P0> : &src_split_launch64<SZ, G, false, P0>,
                            io32};
                e.split = P0;
                v.push_back(e);
            }
        }
    }
}

// The one-wave-per-row source pass (lengths whose lines fit one wave and whose first
// radix is at most 8, so that the four fields' stage-0 inputs fit in registers): G is the
// rows of a block; the LDS holds one line per row and the per-stage twiddle table.  Lengths
// whose last stage splits over P lanes (split_lanes) list that form first.
template <class SZ>
void add_source_wave(std::vector<FastEntry>& v, const char* radices) {
    if constexpr (SZ::TPL <= 64 && SZ::rad(0) <= 8) {
        constexpr int rows = 64 / SZ::TPL;
        // line buffers and twiddles (complex), then the row's wavenumbers (real)
        constexpr size_t cells = (size_t)kSrcWaves * rows * SZ::PITCH + StageTw<SZ>::total() + (SZ::N + 1) / 2;
        constexpr int P = split_lanes<SZ>();
        if constexpr (P > 1) {
            FastEntry e32{SZ::N, kRoleSourceW, kSrcWaves * rows, 64 * kSrcWaves, SZ::TPL, cells * 8, radices,
                          &srcw_launch<SZ, rows, P>, (const void*)fast_source_wave<float, SZ, rows, false, P>, false};
            FastEntry e64{SZ::N, kRoleSourceW, kSrcWaves * rows, 64 * kSrcWaves, SZ::TPL, cells * 16, radices, nullptr,
                          (const void*)fast_source_wave<double, SZ, rows, false, P>, false, 8,
                          &srcw_launch64<SZ, rows, false, P>};
            FastEntry e64io{SZ::N, kRoleSourceW, kSrcWaves * rows, 64 * kSrcWaves, SZ::TPL, cells * 16, radices,
                            nullptr, (const void*)fast_source_wave<double, SZ, rows, true, P>, false, 8,
                            &srcw_launch64<SZ, rows, true, P>, true};
            for (FastEntry* e : {&e32, &e64, &e64io}) {
                e->split = P;
                v.push_back(*e);
            }
        }
        v.push_back({SZ::N, kRoleSourceW, kSrcWaves * rows, 64 * kSrcWaves, SZ::TPL, cells * 8, radices,
                     &fast_launch<SZ, kRoleSourceW, kSrcWaves * rows>,
                     (const void*)fast_source_wave<float, SZ, rows>, false});
        v.push_back({SZ::N, kRoleSourceW, kSrcWaves * rows, 64 * kSrcWaves, SZ::TPL, cells * 16, radices, nullptr,
                     (const void*)fast_source_wave<double, SZ, rows>, false, 8,
                     &fast_launch64<SZ, kRoleSourceW, kSrcWaves * rows>});
        v.push_back({SZ::N, kRoleSourceW, kSrcWaves * rows, 64 * kSrcWaves, SZ::TPL, cells * 16, radices, nullptr,
                     (const void*)fast_source_wave<double, SZ, rows, true>, false, 8,
                     &fast_launch64<SZ, kRoleSourceW, kSrcWaves * rows, true, true>, true});
    }
}

// The default tile width first, then the variants an autotuner may pick: the source
// pass with twice the rows while the block stays within 512 threads (beyond that the
// register budget halves and its accumulators spill), strided passes with half the
// columns.
template <class SZ, int ROLE>
void add_fast_variants(std::vector<FastEntry>& v, const char* radices) {
    constexpr int G = fast_tile(ROLE, SZ::TPL);
    add_fast<SZ, ROLE, G>(v, radices);
    if constexpr (ROLE == kRoleSource) {
        if constexpr (8 * G * SZ::TPL <= 512) add_fast<SZ, ROLE, 2 * G>(v, radices);
    } else if constexpr (ROLE != kRoleContig && G >= 8) {
        add_fast<SZ, ROLE, G / 2>(v, radices);
        // 256-byte rows: fewer, longer HBM bursts per line of the tile
        constexpr bool wide = G == 16 && 32 * SZ::TPL <= 512 && (ROLE == kRoleIntensity || ROLE == kRoleStrided);
        if constexpr (wide) add_fast<SZ, ROLE, 32>(v, radices);
        // the intensity pass without the register prefetch: a second block per CU where the
        // prefetched line was what kept it out
        if constexpr (ROLE == kRoleIntensity) {
            add_fast<SZ, ROLE, G, false>(v, radices);
            if constexpr (wide) add_fast<SZ, ROLE, 32, false>(v, radices);
        }
    }
}

constexpr unsigned kRolesContig = (1u << kRoleSource) | (1u << kRoleContig);
constexpr unsigned kRolesStrided = (1u << kRoleStrided) | (1u << kRoleIntensity) | (1u << kRoleStrided1) |
                                   (1u << kRoleMix1) | (1u << kRoleMix2);
constexpr unsigned kRolesSplitSteps = (1u << kRoleSplit) | (1u << kRoleStridedB) | (1u << kRoleIntensityB) |
                                      (1u << kRoleMix1B) | (1u << kRoleMix2B);

template <class SZ, unsigned ROLES>
void add_fast_roles(std::vector<FastEntry>& v, const char* radices) {
    if constexpr ((ROLES >> kRoleSource) & 1u) {
        add_source_wave<SZ>(v, radices);
        add_source_split<SZ>(v, radices);
        add_fast_variants<SZ, kRoleSource>(v, radices);
    }
    if constexpr ((ROLES >> kRoleContig) & 1u) add_fast_variants<SZ, kRoleContig>(v, radices);
    if constexpr ((ROLES >> kRoleStrided) & 1u) add_fast_variants<SZ, kRoleStrided>(v, radices);
    if constexpr ((ROLES >> kRoleIntensity) & 1u) add_fast_variants<SZ, kRoleIntensity>(v, radices);
    if constexpr ((ROLES >> kRoleStrided1) & 1u)
        add_fast<SZ, kRoleStrided1, fast_tile(kRoleStrided1, SZ::TPL)>(v, radices);
    if constexpr ((ROLES >> kRoleSplit) & 1u) {
        add_split_lds<SZ, fast_tile(kRoleSplit, SZ::TPL), false>(v, radices);
        add_fast<SZ, kRoleSplit, fast_tile(kRoleSplit, SZ::TPL)>(v, radices);
    }
    if constexpr ((ROLES >> kRoleStridedB) & 1u)
        add_fast<SZ, kRoleStridedB, fast_tile(kRoleStridedB, SZ::TPL)>(v, radices);
    if constexpr ((ROLES >> kRoleIntensityB) & 1u)
        add_fast<SZ, kRoleIntensityB, fast_tile(kRoleIntensityB, SZ::TPL)>(v, radices);
    if constexpr ((ROLES >> kRoleMix1) & 1u) {
        add_held<SZ, kRoleMix1, 2>(v, radices);
        add_held<SZ, kRoleMix1, 3>(v, radices);
        add_fast_variants<SZ, kRoleMix1>(v, radices);
    }
    if constexpr ((ROLES >> kRoleMix2) & 1u) {
        add_held<SZ, kRoleMix2, 2>(v, radices);
        add_held<SZ, kRoleMix2, 3>(v, radices);
        add_fast_variants<SZ, kRoleMix2>(v, radices);
    }
    if constexpr ((ROLES >> kRoleSource) & 1u) add_fast64_variants<SZ, kRoleSource>(v, radices);
    if constexpr ((ROLES >> kRoleStrided) & 1u) add_fast64_variants<SZ, kRoleStrided>(v, radices);
    if constexpr ((ROLES >> kRoleIntensity) & 1u) add_fast64_variants<SZ, kRoleIntensity>(v, radices);
    if constexpr ((ROLES >> kRoleMix1) & 1u) add_fast64_variants<SZ, kRoleMix1>(v, radices);
    if constexpr ((ROLES >> kRoleMix2) & 1u) add_fast64_variants<SZ, kRoleMix2>(v, radices);
    // FP64 split steps of the intensity pass of an FP32 plan computing in FP64: the first
    // reads the FP32 spectra and writes an FP64 intermediate (io32), the second reads that
    if constexpr ((ROLES >> kRoleSplit) & 1u) {
        add_split_lds<SZ, fast_tile(kRoleSplit, SZ::TPL) / 2, true>(v, radices);
        add_fast64<SZ, kRoleSplit, fast_tile(kRoleSplit, SZ::TPL) / 2, true, true>(v, radices);
    }
    if constexpr ((ROLES >> kRoleIntensityB) & 1u) {
        add_strided_lds<SZ, kRoleIntensityB, fast_tile(kRoleIntensityB, SZ::TPL) / 2, false>(v, radices);
        add_fast64<SZ, kRoleIntensityB, fast_tile(kRoleIntensityB, SZ::TPL) / 2, true, false>(v, radices);
    }
    if constexpr ((ROLES >> kRoleMix1B) & 1u) {
        add_mix_tabl<SZ, kRoleMix1B, fast_tile(kRoleMix1B, SZ::TPL)>(v, radices);
        add_fast<SZ, kRoleMix1B, fast_tile(kRoleMix1B, SZ::TPL)>(v, radices);
    }
    if constexpr ((ROLES >> kRoleMix2B) & 1u) {
        add_mix_tabl<SZ, kRoleMix2B, fast_tile(kRoleMix2B, SZ::TPL)>(v, radices);
        add_fast<SZ, kRoleMix2B, fast_tile(kRoleMix2B, SZ::TPL)>(v, radices);
    }
}

}  // namespace qpd
