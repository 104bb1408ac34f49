// Host-callable launch wrappers for the CDFCI kernels.  nw = determinant
// words (1 or 2); tab points at the table's SlotW<nw> array.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace mc {

void launch_submatrix(hipStream_t s, int nw, const DevHam& H, DevState* st, const D2* dets, int k,
                      double* Hs, int ldh);
// Subspace update.  xb: the exchange buffer of the solver's worker
// workgroups (kXchgWords words, zeroed once when allocated; used for
// k >= kEigLdsN).
constexpr int kXchgWords = 1152;
void launch_eig(hipStream_t s, DevState* st, int k, int ldh, const double* Hs, const double* px,
                const double* pz, double* dx, double* tabx, int slot_stride, const uint32_t* pslot, int mode,
                uint64_t* xb);
// ept: column elements per thread of the scatter (1, 2, 4, 8 or 16); any value
// gives the same result.  eslot (k columns of vis_words * 64 entries) gets,
// per column element, the slot where the element's probe ended -- its
// determinant's slot, or the empty slot that ended a miss -- from which
// pass 2 resumes the lookup (0xffffffff: probe from the start).
// Line-search update of the single picked coordinate (k = 1).
void launch_ls(hipStream_t s, DevState* st, const double* Hs, const double* px, const double* pz, double* dx,
               double* tabx, int slot_stride, const uint32_t* pslot);
void launch_scatter(hipStream_t s, int nw, const DevHam& H, DevState* st, const D2* pdet, const double* dx,
                    const double* Hs, int ldh, void* tab, uint64_t mask, uint64_t* vis, int vis_words, int k,
                    int ept, long long* ctr, uint32_t* eslot);
// znorm_list (kZNormCap slots): pass 2 lists the slots whose z word B has
// grown past 2^znorm_bits; the merge folds them into word A, and the next
// k_submatrix empties the list.
void launch_pass2(hipStream_t s, int nw, const DevHam& H, DevState* st, const D2* pdet, const double* dx,
                  const double* Hs, int ldh, void* tab, uint64_t mask, const uint64_t* vis, int vis_words, int M,
                  int k, int nchunks, StageEnt* stage, int* stage_cnt, int64_t* part, const uint32_t* eslot,
                  uint32_t* znorm_list);
void launch_merge(hipStream_t s, int nw, DevState* st, const uint32_t* pslot, void* tab, int M, int k,
                  int nchunks, StageEnt* stage, const int* stage_cnt, const int64_t* part, int64_t* col_xz,
                  CandEnt* cand, int* cand_cnt, const double* dx, int count_iteration, long long* ctr,
                  uint64_t nslots, const uint32_t* znorm_list, int merge_tails);
// tree: per-column sorts and a merge tree even where one workgroup could
// sort all k * M candidates (same result either way).
void launch_pick(hipStream_t s, int nw, DevState* st, int k, int M, const CandEnt* cand, const int* cand_cnt,
                 const void* tab, D2* pdet, uint32_t* pslot, double* px, double* pz, PickEnt* pk0, PickEnt* pk1,
                 int* pc0, int* pc1, bool tree);
void launch_init_ref(hipStream_t s, int nw, DevState* st, D2 det, void* tab, uint64_t mask, D2* pdet,
                     uint32_t* pslot, double* px, double* pz, double* dx);
void launch_table_init(hipStream_t s, int nw, void* tab, uint64_t nslots);
// Checkpoint reload: insert distinct (det, x, zA, zB) entries into an empty
// table; re-resolve the candidates' slots.  *fail is set on a full table or
// a candidate missing from it.
void launch_table_insert(hipStream_t s, int nw, void* tab, uint64_t mask, long long n, const uint64_t* det2,
                         const double* x, const int64_t* z2, int* fail);
void launch_cand_resolve(hipStream_t s, int nw, const void* tab, uint64_t mask, CandEnt* cand, const int* cnt,
                         int k, int M, int* fail);
// Random-slot access ceiling (mode 0 insert, 1 update, 2 read); clobbers the table.
void launch_access_probe(hipStream_t s, int nw, void* tab, uint64_t mask, long long n, int mode, uint64_t seed,
                         unsigned long long* sink);
// Column elements of det: two words per determinant in odet.
void launch_column_dump(hipStream_t s, int nw, const DevHam& H, D2 det, double diag, uint64_t* odet, double* oval,
                        unsigned long long* count, long long cap);
// acc[2] (zA, zB fixed point) += sum over dets (x != 0, c = x * scale) and their
// column elements D' of c_D * H_{D'D} (fp64 integrals) * c_D'.
void launch_rayleigh(hipStream_t s, int nw, const DevHam& H, const void* tab, uint64_t mask, const D2* dets,
                     const double* xs, long long n, double scale, int64_t* acc);
// out[3] += {occupied slots, sum of key hashes, sum of entry hashes} (mod 2^64).
void launch_table_digest(hipStream_t s, int nw, const void* tab, uint64_t nslots, unsigned long long* out);
// Occupied slots: two words per determinant in odet, z as (zA, zB) pairs.
void launch_table_compact(hipStream_t s, int nw, const void* tab, uint64_t nslots, uint64_t* odet, double* ox,
                          int64_t* oz, unsigned long long* count, long long cap);

}  // namespace mc
