// Instances of the multi-wave kernel with the older M scores in LDS (scaled
// penalties 2/3/1, i.e. 4/6/2): 16 chunks of 64 diagonals per wave, 4 to 16
// waves per pair (windows of 4096 to 16384 diagonals); offsets doubled for
// sequences of at most 16383 bases; the waves of a pair synchronise with
// their neighbours only.
#include "wfa_regml.h"

namespace tsu {

constexpr int kMlChunks = 16;

template <bool DBL>
static KernelFn pick_regml_u(const Shape& sh) {
    if (sh.bt) {  // CIGAR launches are always queue-fed
        if (!sh.queue) return nullptr;
        switch (sh.nw) {
        case 4: return wfa_regml_kernel<4, 16, true, true, DBL>;
        case 8: return wfa_regml_kernel<8, 16, true, true, DBL>;
        case 16: return wfa_regml_kernel<16, 16, true, true, DBL>;
        default: return nullptr;
        }
    }
    switch (sh.nw) {
    case 4: return sh.queue ? wfa_regml_kernel<4, 16, true, false, DBL> : wfa_regml_kernel<4, 16, false, false, DBL>;
    case 8: return sh.queue ? wfa_regml_kernel<8, 16, true, false, DBL> : wfa_regml_kernel<8, 16, false, false, DBL>;
    case 16: return sh.queue ? wfa_regml_kernel<16, 16, true, false, DBL> : wfa_regml_kernel<16, 16, false, false, DBL>;
    default: return nullptr;
    }
}

KernelFn pick_regml(const Shape& sh) {
    if (!sh.fixed_242 || sh.reg_c != kMlChunks) return nullptr;
    return sh.dbl ? pick_regml_u<true>(sh) : pick_regml_u<false>(sh);
}

int regml_exchange_words(int nw) {
    switch (nw) {
    case 4: return MlLayout<4, kMlChunks>::WORDS;
    case 8: return MlLayout<8, kMlChunks>::WORDS;
    case 16: return MlLayout<16, kMlChunks>::WORDS;
    default: return 0;
    }
}

}  // namespace tsu
