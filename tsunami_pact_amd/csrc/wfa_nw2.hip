// Kernel instances with 2 wave(s) per pair (one translation unit per
// wave count so the variants compile in parallel).
#include "wfa_kernel.h"

namespace tsu {
KernelFn pick_nw2(const Shape& sh) { return pick_nw_impl<2>(sh); }
}  // namespace tsu
