// Single-wave register kernel instances (scaled penalties 2/3/1, i.e. 4/6/2),
// windows of 64, 128, 256 and 512 diagonals: the whole M ring in registers,
// or the older M scores in an int16 LDS ring (sequences up to 16000 bases);
// with records (CIGARs) the int16-ring kernel also takes 1024 diagonals.
#include "wfa_reg.h"
#include "wfa_reg1l.h"

namespace tsu {

KernelFn pick_reg(const Shape& sh) {
    if (!sh.fixed_242) return nullptr;
    if (sh.bt) {  // CIGAR launches: the int16-ring variant with records, queue-fed
        if (!sh.lring || !sh.queue) return nullptr;
        switch (sh.reg_c) {
        case 1: return wfa_reg1l_kernel<1, true, true>;
        case 2: return wfa_reg1l_kernel<2, true, true>;
        case 4: return wfa_reg1l_kernel<4, true, true>;
        case 8: return wfa_reg1l_kernel<8, true, true>;
        case 16: return wfa_reg1l_kernel<16, true, true>;  // CIGARs of windows of 513..1024 diagonals
        default: return nullptr;
        }
    }
    if (sh.lring) {
        switch (sh.reg_c) {
        case 1: return sh.queue ? wfa_reg1l_kernel<1, true> : wfa_reg1l_kernel<1, false>;
        case 2: return sh.queue ? wfa_reg1l_kernel<2, true> : wfa_reg1l_kernel<2, false>;
        case 4: return sh.queue ? wfa_reg1l_kernel<4, true> : wfa_reg1l_kernel<4, false>;
        case 8: return sh.queue ? wfa_reg1l_kernel<8, true> : wfa_reg1l_kernel<8, false>;
        default: return nullptr;
        }
    }
    if (sh.probe) return sh.reg_c == 8 && sh.queue ? wfa_reg_kernel<8, 2, 3, 1, true, true> : nullptr;
    switch (sh.reg_c) {
    case 1: return sh.queue ? wfa_reg_kernel<1, 2, 3, 1, true> : wfa_reg_kernel<1, 2, 3, 1, false>;
    case 2: return sh.queue ? wfa_reg_kernel<2, 2, 3, 1, true> : wfa_reg_kernel<2, 2, 3, 1, false>;
    case 4: return sh.queue ? wfa_reg_kernel<4, 2, 3, 1, true> : wfa_reg_kernel<4, 2, 3, 1, false>;
    case 8: return sh.queue ? wfa_reg_kernel<8, 2, 3, 1, true> : wfa_reg_kernel<8, 2, 3, 1, false>;
    default: return nullptr;
    }
}

}  // namespace tsu
