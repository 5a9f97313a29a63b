// Kernel instances with 8 wave(s) per pair (one translation unit per
// wave count so the variants compile in parallel).
#include "wfa_kernel.h"

namespace tsu {
KernelFn pick_nw8(const Shape& sh) { return pick_nw_impl<8>(sh); }
}  // namespace tsu
