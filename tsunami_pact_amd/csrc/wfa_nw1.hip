// Kernel instances with 1 wave(s) per pair (one translation unit per
// wave count so the variants compile in parallel).
#include "wfa_kernel.h"

namespace tsu {
KernelFn pick_nw1(const Shape& sh) { return pick_nw_impl<1>(sh); }
}  // namespace tsu
