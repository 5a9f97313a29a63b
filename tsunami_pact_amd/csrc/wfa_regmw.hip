// Multi-wave register-resident kernel instances (scaled penalties 2/3/1, i.e.
// 4/6/2): 8 chunks of 64 diagonals per wave, 2 to 16 waves per pair
// (windows of 1024 to 8192 diagonals).
#include "wfa_regmw.h"

namespace tsu {

constexpr int kMwChunks = 8;

KernelFn pick_regmw(const Shape& sh) {
    if (!sh.fixed_242 || sh.bt || sh.reg_c != kMwChunks) return nullptr;
    switch (sh.nw) {
    case 2: return sh.queue ? wfa_regmw_kernel<2, 8, 2, 3, 1, true> : wfa_regmw_kernel<2, 8, 2, 3, 1, false>;
    case 4: return sh.queue ? wfa_regmw_kernel<4, 8, 2, 3, 1, true> : wfa_regmw_kernel<4, 8, 2, 3, 1, false>;
    case 8: return sh.queue ? wfa_regmw_kernel<8, 8, 2, 3, 1, true> : wfa_regmw_kernel<8, 8, 2, 3, 1, false>;
    case 16: return sh.queue ? wfa_regmw_kernel<16, 8, 2, 3, 1, true> : wfa_regmw_kernel<16, 8, 2, 3, 1, false>;
    default: return nullptr;
    }
}

int regmw_exchange_words(int nw) {
    switch (nw) {
    case 2: return MwLayout<2, kMwChunks, 5>::WORDS;
    case 4: return MwLayout<4, kMwChunks, 5>::WORDS;
    case 8: return MwLayout<8, kMwChunks, 5>::WORDS;
    case 16: return MwLayout<16, kMwChunks, 5>::WORDS;
    default: return 0;
    }
}

}  // namespace tsu
