// Host-callable launch wrappers for the energy-gradient kernels (nw =
// determinant words; tab points at the table's SlotW<nw> array).
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace mc {

void launch_compact_x(hipStream_t s, int nw, const void* tab, uint64_t nslots, D2* odet, double* ox,
                      unsigned long long* count, long long cap);
void launch_sum_sq(hipStream_t s, const double* x, long long n, double scale, int64_t* acc);
void launch_energy_grad(hipStream_t s, int nw, const DevHam& H, const void* tab, uint64_t mask, const D2* dets,
                        const double* xs, long long n, double scale, int64_t* acc_h, int64_t* acc_e);

}  // namespace mc
