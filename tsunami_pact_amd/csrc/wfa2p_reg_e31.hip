// Instances of the 2-piece register kernel for gap extensions e1 = 3,
// e2 = 1 after the gcd (one translation unit per pair of extensions).
#include "wfa2p_reg.h"

namespace tsu {
KernelFn pick_2p_reg_31(const Shape& sh) { return pick_2p_reg_impl<3, 1>(sh); }
}  // namespace tsu
