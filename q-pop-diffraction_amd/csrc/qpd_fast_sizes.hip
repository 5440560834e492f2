// The transform lengths with compile-time kernels, and their radix sequences.
//
// Compiled kFastParts times with -DQPD_FAST_PART=p; part p instantiates the entries
// whose position in the list is p modulo kFastParts, so the parts build in parallel.
// The first radix is the stage that reads its inputs (global memory in the strided
// passes), the last the one whose outputs are accumulated.  Contiguous passes keep
// radix 2 out of long lengths, where it would set the threads per line (N / R_min);
// strided passes take the largest radix first, which leaves the fewest accumulators
// (NOUT x R_last) per thread.  Lengths that can be a step of a split long transform
// also get the split roles; lengths above 512 only get the contiguous roles (long
// strided axes are split).
#include "qpd_fast_register.hpp"

#ifndef QPD_FAST_PART
#error "compile with -DQPD_FAST_PART=<part>"
#endif

#define QPD_CAT2(a, b) a##b
#define QPD_CAT(a, b) QPD_CAT2(a, b)

namespace qpd {

namespace {
template <int I, class SZ, unsigned ROLES>
void add_if_mine(std::vector<FastEntry>& v, const char* radices) {
    if constexpr (I % kFastParts == QPD_FAST_PART) add_fast_roles<SZ, ROLES>(v, radices);
}
}  // namespace

void QPD_CAT(register_fast_part, QPD_FAST_PART)(std::vector<FastEntry>& v) {
    add_if_mine<0, Sizes<16, 4, 4>, kRolesContig>(v, "4x4");
    add_if_mine<1, Sizes<20, 4, 5>, kRolesContig>(v, "4x5");
    add_if_mine<2, Sizes<24, 8, 3>, kRolesContig>(v, "8x3");
    add_if_mine<3, Sizes<25, 5, 5>, kRolesContig>(v, "5x5");
    add_if_mine<4, Sizes<28, 4, 7>, kRolesContig>(v, "4x7");
    add_if_mine<5, Sizes<30, 2, 3, 5>, kRolesContig>(v, "2x3x5");
    add_if_mine<6, Sizes<32, 8, 4>, kRolesContig>(v, "8x4");
    add_if_mine<7, Sizes<34, 2, 17>, kRolesContig>(v, "2x17");
    add_if_mine<8, Sizes<36, 4, 3, 3>, kRolesContig>(v, "4x3x3");
    add_if_mine<9, Sizes<40, 8, 5>, kRolesContig>(v, "8x5");
    add_if_mine<10, Sizes<42, 2, 3, 7>, kRolesContig>(v, "2x3x7");
    add_if_mine<11, Sizes<44, 4, 11>, kRolesContig>(v, "4x11");
    add_if_mine<12, Sizes<48, 4, 4, 3>, kRolesContig>(v, "4x4x3");
    add_if_mine<13, Sizes<50, 2, 5, 5>, kRolesContig>(v, "2x5x5");
    add_if_mine<14, Sizes<52, 4, 13>, kRolesContig>(v, "4x13");
    add_if_mine<15, Sizes<56, 8, 7>, kRolesContig>(v, "8x7");
    add_if_mine<16, Sizes<60, 4, 3, 5>, kRolesContig>(v, "4x3x5");
    add_if_mine<17, Sizes<64, 8, 8>, kRolesContig>(v, "8x8");
    add_if_mine<18, Sizes<68, 4, 17>, kRolesContig>(v, "4x17");
    add_if_mine<19, Sizes<72, 8, 3, 3>, kRolesContig>(v, "8x3x3");
    add_if_mine<20, Sizes<80, 4, 4, 5>, kRolesContig>(v, "4x4x5");
    add_if_mine<21, Sizes<84, 4, 3, 7>, kRolesContig>(v, "4x3x7");
    add_if_mine<22, Sizes<88, 8, 11>, kRolesContig>(v, "8x11");
    add_if_mine<23, Sizes<96, 8, 4, 3>, kRolesContig>(v, "8x4x3");
    add_if_mine<24, Sizes<100, 4, 5, 5>, kRolesContig>(v, "4x5x5");
    add_if_mine<25, Sizes<104, 8, 13>, kRolesContig>(v, "8x13");
    add_if_mine<26, Sizes<112, 4, 4, 7>, kRolesContig>(v, "4x4x7");
    add_if_mine<27, Sizes<120, 8, 3, 5>, kRolesContig>(v, "8x3x5");
    add_if_mine<28, Sizes<125, 5, 5, 5>, kRolesContig>(v, "5x5x5");
    add_if_mine<29, Sizes<128, 4, 4, 8>, kRolesContig>(v, "4x4x8");
    add_if_mine<30, Sizes<136, 8, 17>, kRolesContig>(v, "8x17");
    add_if_mine<31, Sizes<144, 4, 4, 3, 3>, kRolesContig>(v, "4x4x3x3");
    add_if_mine<32, Sizes<160, 4, 8, 5>, kRolesContig>(v, "4x8x5");
    add_if_mine<33, Sizes<176, 4, 4, 11>, kRolesContig>(v, "4x4x11");
    add_if_mine<34, Sizes<192, 8, 8, 3>, kRolesContig>(v, "8x8x3");
    add_if_mine<35, Sizes<200, 8, 5, 5>, kRolesContig>(v, "8x5x5");
    add_if_mine<36, Sizes<208, 4, 4, 13>, kRolesContig>(v, "4x4x13");
    add_if_mine<37, Sizes<224, 8, 4, 7>, kRolesContig>(v, "8x4x7");
    add_if_mine<38, Sizes<240, 4, 4, 3, 5>, kRolesContig>(v, "4x4x3x5");
    add_if_mine<39, Sizes<256, 4, 8, 8>, kRolesContig>(v, "4x8x8");
    add_if_mine<112, Sizes<288, 8, 6, 6>, kRolesContig>(v, "8x6x6");  // 48 threads a line
    add_if_mine<40, Sizes<288, 8, 4, 3, 3>, kRolesContig>(v, "8x4x3x3");
    add_if_mine<41, Sizes<320, 8, 8, 5>, kRolesContig>(v, "8x8x5");
    add_if_mine<114, Sizes<384, 8, 8, 6>, kRolesContig>(v, "8x8x6");  // 64 threads a line
    add_if_mine<42, Sizes<384, 4, 4, 8, 3>, kRolesContig>(v, "4x4x8x3");
    add_if_mine<43, Sizes<400, 4, 4, 5, 5>, kRolesContig>(v, "4x4x5x5");
    add_if_mine<44, Sizes<448, 8, 8, 7>, kRolesContig>(v, "8x8x7");
    add_if_mine<45, Sizes<480, 8, 4, 3, 5>, kRolesContig>(v, "8x4x3x5");
    add_if_mine<46, Sizes<500, 4, 5, 5, 5>, kRolesContig>(v, "4x5x5x5");
    add_if_mine<47, Sizes<512, 8, 8, 8>, kRolesContig>(v, "8x8x8");
    add_if_mine<48, Sizes<640, 4, 4, 8, 5>, kRolesContig>(v, "4x4x8x5");
    add_if_mine<49, Sizes<768, 8, 8, 4, 3>, kRolesContig>(v, "8x8x4x3");
    add_if_mine<50, Sizes<800, 8, 4, 5, 5>, kRolesContig>(v, "8x4x5x5");
    add_if_mine<51, Sizes<1000, 8, 5, 5, 5>, kRolesContig>(v, "8x5x5x5");
    add_if_mine<52, Sizes<1024, 8, 8, 4, 4>, kRolesContig>(v, "8x8x4x4");
    add_if_mine<53, Sizes<16, 4, 4>, kRolesStrided | kRolesSplitSteps>(v, "4x4");
    add_if_mine<54, Sizes<20, 5, 4>, kRolesStrided | kRolesSplitSteps>(v, "5x4");
    add_if_mine<55, Sizes<24, 8, 3>, kRolesStrided | kRolesSplitSteps>(v, "8x3");
    add_if_mine<56, Sizes<25, 5, 5>, kRolesStrided | kRolesSplitSteps>(v, "5x5");
    add_if_mine<57, Sizes<28, 7, 4>, kRolesStrided>(v, "7x4");
    add_if_mine<58, Sizes<30, 5, 3, 2>, kRolesStrided>(v, "5x3x2");
    add_if_mine<59, Sizes<32, 8, 4>, kRolesStrided | kRolesSplitSteps>(v, "8x4");
    add_if_mine<60, Sizes<34, 17, 2>, kRolesStrided>(v, "17x2");
    add_if_mine<61, Sizes<36, 4, 3, 3>, kRolesStrided>(v, "4x3x3");
    add_if_mine<62, Sizes<40, 8, 5>, kRolesStrided | kRolesSplitSteps>(v, "8x5");
    add_if_mine<63, Sizes<42, 7, 3, 2>, kRolesStrided>(v, "7x3x2");
    add_if_mine<64, Sizes<44, 11, 4>, kRolesStrided>(v, "11x4");
    add_if_mine<65, Sizes<48, 4, 4, 3>, kRolesStrided | kRolesSplitSteps>(v, "4x4x3");
    add_if_mine<66, Sizes<50, 5, 5, 2>, kRolesStrided | kRolesSplitSteps>(v, "5x5x2");
    add_if_mine<67, Sizes<52, 13, 4>, kRolesStrided>(v, "13x4");
    add_if_mine<68, Sizes<56, 8, 7>, kRolesStrided>(v, "8x7");
    add_if_mine<69, Sizes<60, 5, 4, 3>, kRolesStrided>(v, "5x4x3");
    add_if_mine<70, Sizes<64, 8, 8>, kRolesStrided | kRolesSplitSteps>(v, "8x8");
    add_if_mine<71, Sizes<68, 17, 4>, kRolesStrided>(v, "17x4");
    add_if_mine<72, Sizes<72, 8, 3, 3>, kRolesStrided>(v, "8x3x3");
    add_if_mine<73, Sizes<80, 5, 4, 4>, kRolesStrided | kRolesSplitSteps>(v, "5x4x4");
    add_if_mine<74, Sizes<84, 7, 4, 3>, kRolesStrided>(v, "7x4x3");
    add_if_mine<75, Sizes<88, 11, 8>, kRolesStrided>(v, "11x8");
    add_if_mine<76, Sizes<96, 8, 4, 3>, kRolesStrided | kRolesSplitSteps>(v, "8x4x3");
    add_if_mine<77, Sizes<100, 5, 5, 4>, kRolesStrided | kRolesSplitSteps>(v, "5x5x4");
    add_if_mine<78, Sizes<104, 13, 8>, kRolesStrided>(v, "13x8");
    add_if_mine<79, Sizes<112, 7, 4, 4>, kRolesStrided>(v, "7x4x4");
    add_if_mine<80, Sizes<120, 8, 5, 3>, kRolesStrided>(v, "8x5x3");
    add_if_mine<81, Sizes<125, 5, 5, 5>, kRolesStrided | kRolesSplitSteps>(v, "5x5x5");
    add_if_mine<82, Sizes<128, 8, 4, 4>, kRolesStrided | kRolesSplitSteps>(v, "8x4x4");
    add_if_mine<83, Sizes<136, 17, 8>, kRolesStrided>(v, "17x8");
    add_if_mine<84, Sizes<144, 4, 4, 3, 3>, kRolesStrided>(v, "4x4x3x3");
    add_if_mine<85, Sizes<160, 8, 5, 4>, kRolesStrided>(v, "8x5x4");
    add_if_mine<86, Sizes<176, 11, 4, 4>, kRolesStrided>(v, "11x4x4");
    add_if_mine<87, Sizes<192, 8, 8, 3>, kRolesStrided>(v, "8x8x3");
    add_if_mine<88, Sizes<200, 8, 5, 5>, kRolesStrided>(v, "8x5x5");
    add_if_mine<89, Sizes<208, 13, 4, 4>, kRolesStrided>(v, "13x4x4");
    add_if_mine<90, Sizes<224, 8, 7, 4>, kRolesStrided>(v, "8x7x4");
    add_if_mine<91, Sizes<240, 5, 4, 4, 3>, kRolesStrided>(v, "5x4x4x3");
    add_if_mine<92, Sizes<256, 8, 8, 4>, kRolesStrided>(v, "8x8x4");
    add_if_mine<109, Sizes<288, 8, 6, 6>, kRolesStrided>(v, "8x6x6");
    add_if_mine<93, Sizes<288, 8, 4, 3, 3>, kRolesStrided>(v, "8x4x3x3");
    add_if_mine<94, Sizes<320, 8, 8, 5>, kRolesStrided>(v, "8x8x5");
    add_if_mine<111, Sizes<384, 8, 8, 6>, kRolesStrided>(v, "8x8x6");
    add_if_mine<95, Sizes<384, 8, 4, 4, 3>, kRolesStrided>(v, "8x4x4x3");
    add_if_mine<96, Sizes<400, 5, 5, 4, 4>, kRolesStrided>(v, "5x5x4x4");
    add_if_mine<97, Sizes<448, 8, 8, 7>, kRolesStrided>(v, "8x8x7");
    add_if_mine<98, Sizes<480, 8, 5, 4, 3>, kRolesStrided>(v, "8x5x4x3");
    add_if_mine<99, Sizes<500, 5, 5, 5, 4>, kRolesStrided>(v, "5x5x5x4");
    add_if_mine<100, Sizes<512, 8, 8, 8>, kRolesStrided>(v, "8x8x8");
    // large radix first: fewer accumulators in the source pass, for the autotuner to weigh
    add_if_mine<101, Sizes<68, 17, 4>, kRolesContig>(v, "17x4");
    add_if_mine<102, Sizes<34, 17, 2>, kRolesContig>(v, "17x2");
    add_if_mine<103, Sizes<136, 17, 8>, kRolesContig>(v, "17x8");
    add_if_mine<104, Sizes<52, 13, 4>, kRolesContig>(v, "13x4");
    add_if_mine<105, Sizes<104, 13, 8>, kRolesContig>(v, "13x8");
    add_if_mine<106, Sizes<44, 11, 4>, kRolesContig>(v, "11x4");
    add_if_mine<107, Sizes<88, 11, 8>, kRolesContig>(v, "11x8");
}

}  // namespace qpd
