// Host-callable launch wrappers for the energy-gradient kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace mc {

void launch_compact_x(hipStream_t s, const Slot* tab, uint64_t nslots, uint64_t* odet, double* ox,
                      unsigned long long* count, long long cap);
void launch_sum_sq(hipStream_t s, const double* x, long long n, double scale, int64_t* acc);
void launch_energy_grad(hipStream_t s, const DevHam& H, const Slot* tab, uint64_t mask, const uint64_t* dets,
                        const double* xs, long long n, double scale, int64_t* acc_h, int64_t* acc_e);

}  // namespace mc
