// Register-resident kernel instances (scaled penalties 2/3/1, i.e. 4/6/2),
// windows of 64, 128, 256 and 512 diagonals.
#include "wfa_reg.h"

namespace tsu {

KernelFn pick_reg(const Shape& sh) {
    if (!sh.fixed_242 || sh.bt) return nullptr;
    switch (sh.reg_c) {
    case 1: return sh.queue ? wfa_reg_kernel<1, 2, 3, 1, true> : wfa_reg_kernel<1, 2, 3, 1, false>;
    case 2: return sh.queue ? wfa_reg_kernel<2, 2, 3, 1, true> : wfa_reg_kernel<2, 2, 3, 1, false>;
    case 4: return sh.queue ? wfa_reg_kernel<4, 2, 3, 1, true> : wfa_reg_kernel<4, 2, 3, 1, false>;
    case 8: return sh.queue ? wfa_reg_kernel<8, 2, 3, 1, true> : wfa_reg_kernel<8, 2, 3, 1, false>;
    default: return nullptr;
    }
}

}  // namespace tsu
