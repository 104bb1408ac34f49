// Host-callable launch wrappers for the CDFCI kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace mc {

void launch_submatrix(hipStream_t s, const DevHam& H, const DevState* st, const uint64_t* dets, int k,
                      double* Hs, int ldh);
void launch_eig(hipStream_t s, DevState* st, int k, int ldh, const double* Hs, const double* px,
                const double* pz, double* dx, Slot* tab, const uint32_t* pslot, int mode);
void launch_scatter(hipStream_t s, const DevHam& H, DevState* st, const uint64_t* pdet, const double* dx,
                    const double* Hs, int ldh, Slot* tab, uint64_t mask, uint64_t* vis, int vis_words,
                    int k, int nchunks);
void launch_pass2(hipStream_t s, const DevHam& H, DevState* st, const uint64_t* pdet, const double* dx,
                  const double* Hs, int ldh, Slot* tab, uint64_t mask, const uint64_t* vis, int vis_words,
                  int M, int k, int nchunks, StageEnt* stage, int* stage_cnt, int64_t* part);
void launch_merge(hipStream_t s, DevState* st, const uint64_t* pdet, const uint32_t* pslot, Slot* tab,
                  int M, int k, int nchunks, const StageEnt* stage, const int* stage_cnt,
                  const int64_t* part, int64_t* col_xz, CandEnt* cand, int* cand_cnt, const double* dx,
                  int count_iteration);
void launch_pick(hipStream_t s, DevState* st, int k, int M, const CandEnt* cand, const int* cand_cnt,
                 const Slot* tab, uint64_t* pdet, uint32_t* pslot, double* px, double* pz, PickEnt* pk0,
                 PickEnt* pk1, int* pc0, int* pc1);
void launch_init_ref(hipStream_t s, DevState* st, uint64_t det, Slot* tab, uint64_t mask, uint64_t* pdet,
                     uint32_t* pslot, double* px, double* pz, double* dx);
void launch_column_dump(hipStream_t s, const DevHam& H, uint64_t det, double diag, uint64_t* odet,
                        double* oval, unsigned long long* count, long long cap);
void launch_table_compact(hipStream_t s, const Slot* tab, uint64_t nslots, uint64_t* odet, double* ox,
                          int64_t* oz, unsigned long long* count, long long cap);

}  // namespace mc
