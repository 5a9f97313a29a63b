// cpu_check.h -- score-only CPU WFA behind `wfa_gpu ... check=1`.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/tsunami_wfa.h"

// Scores of every pair of ds (gap-affine, penalties x/o/e).  Returns the
// number of pairs that failed (score slot -1).  threads_used receives the
// OpenMP team size.
int cpu_check_scores(const tsu_dataset* ds, int32_t x, int32_t o, int32_t e, std::vector<int32_t>& out,
                     int* threads_used);
