// Host engine behind the C ABI of include/mcdfci.h: builds the device
// integral tables, owns the HBM wavefunction table and iteration buffers,
// and sequences the kernels of one CDFCI iteration on one HIP stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mcdfci.h"
#include "device.h"
#include "fcidump.h"
#include "grad.h"
#include "kernels.h"

using namespace mc;

// Scratch of mcdfci_eig_update_noalloc, the subspace update that calls no
// device allocator: one copy per GPU (each device holds its own at its own
// address), kept until the process ends, sized for k <= kMaxK:
//   [DevState, 32 doubles] [H_sub k*k] [x k] [z k] [dx k]  (packed by k)
//   and, at a fixed place after the largest of those, the subspace solver's
//   exchange buffer (kXchgWords words; zero at load like the whole array).
// Host code never passes `scratch` itself (a host-side stand-in) to a
// kernel or a copy: it uses the per-device address from
// hipGetSymbolAddress, taken once per device at mcdfci_create.
constexpr size_t kNoAllocState = 32;
constexpr size_t kNoAllocXchg = kNoAllocState + (size_t)kMaxK * kMaxK + 3 * (size_t)kMaxK;
constexpr size_t kNoAllocDoubles = kNoAllocXchg + (size_t)kXchgWords;
static_assert(sizeof(DevState) <= kNoAllocState * sizeof(double), "DevState outgrows its scratch slot");
static __device__ __align__(256) double scratch[kNoAllocDoubles];


namespace {

std::string g_err;

std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_addr;  // device ordinal -> address of its `scratch`

// Mirrors the upper triangle (j > i) of the row-major k x k matrix into the
// lower one: the subspace update reads both triangles.
__global__ void k_mirror_upper(double* H, int k) {
  const long long n = (long long)k * k;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / k), j = (int)(e - (long long)i * k);
    if (j > i) H[(size_t)j * k + i] = H[(size_t)i * k + j];
  }
}

int binom2(int n) { return n < 2 ? 0 : n * (n - 1) / 2; }

}  // namespace

struct mcdfci_ctx {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  SpatialHam ham;
  mcdfci_opts opts{};
  DevHam dh{};
  // integral tables (one allocation each)
  double* d_eri = nullptr;
  float* d_eri32 = nullptr;
  double* d_h1 = nullptr;
  float* d_h132 = nullptr;
  double* d_ss = nullptr;
  double* d_so = nullptr;
  float* d_ss32 = nullptr;
  float* d_so32 = nullptr;
  uint8_t* d_sok = nullptr;
  double* d_d1 = nullptr;
  double* d_d2 = nullptr;
  // wavefunction table: nslots entries of SlotW<nw>
  void* tab = nullptr;
  int nw = 1;    // determinant words on the device (1: up to 31 orbitals, 2: up to 63)
  int nio = 1;   // words per determinant in the C ABI arrays (src/run_wrapper.h:181)
  uint64_t nslots = 0;
  double max_size = 0.0;
  int table_pow = 0;
  // solver state and iteration buffers
  DevState* st = nullptr;
  DevState* st_scratch = nullptr;
  double* noalloc = nullptr;  // this device's copy of `scratch` (mcdfci_eig_update_noalloc)
  int k = 1, M = 1;
  D2* pdet = nullptr;
  uint32_t* pslot = nullptr;
  double* px = nullptr;
  double* pz = nullptr;
  double* dx = nullptr;
  double* Hs = nullptr;
  uint64_t* vis = nullptr;
  int vis_words = 0;
  int nchunks = 0;
  StageEnt* stage = nullptr;
  int* stage_cnt = nullptr;
  int64_t* part = nullptr;
  int64_t* col_xz = nullptr;
  CandEnt* cand = nullptr;
  int* cand_cnt = nullptr;
  PickEnt* pk0 = nullptr;  // pick lists for k * M above one workgroup's sort
  PickEnt* pk1 = nullptr;
  int* pc0 = nullptr;
  int* pc1 = nullptr;
  long long* ctr = nullptr;  // striped iteration counters
  uint32_t* eslot = nullptr;  // per column element: the slot where the scatter's probe ended
  uint32_t* znorm = nullptr;  // slots whose z word B pass 2 found near its range (kZNormCap)
  uint64_t* xbuf = nullptr;   // exchange buffer of the subspace solver's workgroups (kXchgWords)
  int znorm_bits = 61;        // fold threshold (MCDFCI_ZNORM_BITS lowers it to exercise the fold)
  int merge_tails = 7;        // chunk lists whose tails the column merge takes in LDS (MCDFCI_MERGE_TAILS
                              // lowers it to exercise the full tournament it falls back to)
  size_t iter_bytes = 0;
  // launch choices (result-neutral; set by mcdfci_autotune)
  int scatter_ept = 4;
  bool pick_tree = false;
  // captured iterations (mcdfci_set_graph)
  int graph_iters = 0;
  hipGraphExec_t graph = nullptr;
  bool ready = false;
  bool timing = false;
  double timers[6] = {0, 0, 0, 0, 0, 0};
  double ktimes[6] = {0, 0, 0, 0, 0, 0};  // pick, submatrix, eig, scatter, pass2, merge
  long long ktimed = 0;
  hipEvent_t ev[7] = {};
  std::vector<void*> allocs;
};

namespace {

int fail(mcdfci_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_err = msg;
  return code;
}

#define HIPCHECK(c, x)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) return fail((c), MCDFCI_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(mcdfci_ctx* c, T** p, size_t count) {
  void* q = nullptr;
  size_t bytes = std::max<size_t>(count * sizeof(T), 16);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, MCDFCI_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                     hipGetErrorString(e));
  }
  c->allocs.push_back(q);
  *p = (T*)q;
  return MCDFCI_OK;
}

void dfree(mcdfci_ctx* c, void* p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

template <class T>
int upload(mcdfci_ctx* c, T** p, const std::vector<T>& v) {
  int rc = dalloc(c, p, v.size());
  if (rc) return rc;
  HIPCHECK(c, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return MCDFCI_OK;
}

// Enumeration layout of the (na, nb) sector.
void set_sector(DevHam& H, int na, int nb) {
  H.na = na;
  H.nb = nb;
  H.nva = H.M - na;
  H.nvb = H.M - nb;
  H.ne = na + nb;
  H.cpa_o = binom2(na);
  H.cpb_o = binom2(nb);
  H.cpa_v = binom2(H.nva);
  H.cpb_v = binom2(H.nvb);
  H.n_sa = na * H.nva;
  H.n_sb = nb * H.nvb;
  H.n_daa = H.cpa_o * H.cpa_v;
  H.n_dbb = H.cpb_o * H.cpb_v;
  H.n_dab = na * nb * H.nva * H.nvb;
  H.off_sb = 1 + H.n_sa;
  H.off_daa = H.off_sb + H.n_sb;
  H.off_dbb = H.off_daa + H.n_daa;
  H.off_dab = H.off_dbb + H.n_dbb;
  H.n_enum = H.off_dab + H.n_dab;
}

void det_sector(const D2& det, int& na, int& nb) {
  na = __builtin_popcountll(det.lo & 0x5555555555555555ULL) + __builtin_popcountll(det.hi & 0x5555555555555555ULL);
  nb = __builtin_popcountll(det.lo & 0xAAAAAAAAAAAAAAAAULL) + __builtin_popcountll(det.hi & 0xAAAAAAAAAAAAAAAAULL);
}

D2 load_det(const uint64_t* p, int nio) { return D2{p[0], nio > 1 ? p[1] : 0}; }

void store_det(uint64_t* p, const D2& d, int nio) {
  p[0] = d.lo;
  if (nio > 1) p[1] = d.hi;
}

bool det_less(const D2& a, const D2& b) { return a.hi != b.hi ? a.hi < b.hi : a.lo < b.lo; }

size_t slot_bytes(int nw) { return nw == 1 ? sizeof(SlotW<1>) : sizeof(SlotW<2>); }

// Checks a determinant against the device limits of its word count: no bits
// beyond 2 * norb spin-orbitals, and for two words a sector whose alpha or
// beta count is below 32 (an empty slot's word 0 holds all ones), with 1 to
// 94 electrons (the empty determinant keys a reserved slot).
int check_det(mcdfci_ctx* c, const D2& det, const char* what) {
  const int ns = 2 * c->ham.norb;
  const bool beyond = ns < 64 ? ((det.lo >> ns) != 0 || det.hi != 0) : ns < 128 && (det.hi >> (ns - 64)) != 0;
  if (beyond) return fail(c, MCDFCI_E_INPUT, std::string(what) + " has bits beyond NORB");
  int na, nb;
  det_sector(det, na, nb);
  if (c->nw == 2 && (std::min(na, nb) >= 32 || na + nb > 94 || na + nb < 1))
    return fail(c, MCDFCI_E_UNSUPPORTED,
                "two-word determinants need 1 to 94 electrons with fewer than 32 alpha or beta electrons");
  return MCDFCI_OK;
}

// The fp64 tables derived from the integrals: single-excitation sums
// S_same/S_opp [p][r][k], the single-excitation filter [p][r] (one block
// per spin when the integrals are unrestricted, else one shared block) and
// the per-spin-orbital diagonal terms.  For an electron of spin s moving
// p -> r past an occupied orbital k (src/hamiltonian.h:176-219):
//   S_same[s][p][r][k] = (p r|k k)_ss - (p k|k r)_ss      (k of spin s)
//   S_opp[s][p][r][k]  = (p r|k k)_{s s'}                 (k of the other spin)
// with (p r|k k)_{ab} = eri_ab[p][r][k][k] and (p r|k k)_{ba} = eri_ab[k][k][p][r].
struct Fp64Tables {
  std::vector<double> ss, so, d1, d2;
  std::vector<uint8_t> sok;
};

Fp64Tables make_tables(const SpatialHam& S, double threshold) {
  const int M = S.norb, ns = 2 * M;
  const size_t M2 = (size_t)M * M, M3 = M2 * M;
  const int nsp = S.uhf ? 2 : 1;
  auto X = [&](const double* e, int p, int q, int r, int s) { return e[(((size_t)p * M + q) * M + r) * M + s]; };
  Fp64Tables T;
  T.ss.assign(nsp * M3, 0.0);
  T.so.assign(nsp * M3, 0.0);
  T.d1.assign(ns, 0.0);
  T.d2.assign((size_t)ns * ns, 0.0);
  T.sok.assign(nsp * M2, 0);
  const double* eab = S.eri_block(2);
  for (int sp = 0; sp < nsp; ++sp) {
    const double* hs = S.h_spin(sp);
    const double* es = S.eri_block(sp);
    for (int p = 0; p < M; ++p)
      for (int r = 0; r < M; ++r) {
        double mx = std::fabs(hs[p * M + r]);
        for (int k = 0; k < M; ++k) {
          double a = X(es, p, r, k, k);
          double b = X(es, p, k, k, r);
          double o = !S.uhf ? a : (sp == 0 ? X(eab, p, r, k, k) : X(eab, k, k, p, r));
          T.ss[sp * M3 + ((size_t)p * M + r) * M + k] = a - b;
          T.so[sp * M3 + ((size_t)p * M + r) * M + k] = o;
          mx = std::max(mx, std::fabs(a - b));
          mx = std::max(mx, std::fabs(o));
        }
        T.sok[sp * M2 + p * M + r] = mx > threshold ? 1 : 0;
      }
  }
  for (int i = 0; i < ns; ++i) T.d1[i] = S.h_spin(i & 1)[(i >> 1) * M + (i >> 1)];
  for (int i = 0; i < ns; ++i)
    for (int j = i + 1; j < ns; ++j) {
      int p = i >> 1, q = j >> 1;
      double v;
      if (!S.uhf) {
        v = X(S.eri.data(), p, p, q, q);
        if ((i & 1) == (j & 1)) v = v - X(S.eri.data(), p, q, q, p);
      } else if ((i & 1) == (j & 1)) {
        const double* es = S.eri_block(i & 1);
        v = X(es, p, p, q, q) - X(es, p, q, q, p);
      } else {
        v = (i & 1) == 0 ? X(eab, p, p, q, q) : X(eab, q, q, p, p);
      }
      T.d2[(size_t)i * ns + j] = v;
    }
  return T;
}

// The integral arrays as the device holds them: the alpha-alpha,
// beta-beta and alpha-beta ERI blocks one after another (one block when
// the integrals are restricted), and the alpha and beta h.
void device_integrals(const SpatialHam& S, std::vector<double>& eri, std::vector<double>& h) {
  eri = S.eri;
  h = S.h;
  if (S.uhf) {
    eri.insert(eri.end(), S.eri_bb.begin(), S.eri_bb.end());
    eri.insert(eri.end(), S.eri_ab.begin(), S.eri_ab.end());
    h.insert(h.end(), S.hb.begin(), S.hb.end());
  }
}

// Block strides of the device tables: zero for restricted integrals, so
// both spins read the one block.
void set_strides(DevHam& H, const SpatialHam& S) {
  const size_t M2 = (size_t)S.norb * S.norb, M3 = M2 * S.norb, M4 = M2 * M2;
  H.h_so = S.uhf ? M2 : 0;
  H.s_so = S.uhf ? M3 : 0;
  H.e_bb = S.uhf ? M4 : 0;
  H.e_ab = S.uhf ? 2 * M4 : 0;
}

int build_tables(mcdfci_ctx* c) {
  const SpatialHam& S = c->ham;
  const int M = S.norb, ns = 2 * M;
  const Fp64Tables T = make_tables(S, c->opts.threshold);
  std::vector<double> eri, h1;
  device_integrals(S, eri, h1);
  std::vector<float> eri32(eri.size()), h132(h1.size()), ss32(T.ss.size()), so32(T.so.size());
  for (size_t i = 0; i < eri.size(); ++i) eri32[i] = (float)eri[i];
  for (size_t i = 0; i < h1.size(); ++i) h132[i] = (float)h1[i];
  for (size_t i = 0; i < T.ss.size(); ++i) {
    ss32[i] = (float)T.ss[i];
    so32[i] = (float)T.so[i];
  }
  int rc;
  if ((rc = upload(c, &c->d_eri, eri))) return rc;
  if ((rc = upload(c, &c->d_eri32, eri32))) return rc;
  if ((rc = upload(c, &c->d_h1, h1))) return rc;
  if ((rc = upload(c, &c->d_h132, h132))) return rc;
  if ((rc = upload(c, &c->d_ss, T.ss))) return rc;
  if ((rc = upload(c, &c->d_so, T.so))) return rc;
  if ((rc = upload(c, &c->d_ss32, ss32))) return rc;
  if ((rc = upload(c, &c->d_so32, so32))) return rc;
  if ((rc = upload(c, &c->d_sok, T.sok))) return rc;
  if ((rc = upload(c, &c->d_d1, T.d1))) return rc;
  if ((rc = upload(c, &c->d_d2, T.d2))) return rc;
  DevHam& H = c->dh;
  H.M = M;
  H.ns = ns;
  H.mixed = c->opts.precision == MCDFCI_PREC_MIXED ? 1 : 0;
  H.core = S.core;
  H.thr = c->opts.threshold;
  set_strides(H, S);
  H.eri = c->d_eri;
  H.eri32 = c->d_eri32;
  H.h1 = c->d_h1;
  H.h132 = c->d_h132;
  H.s_same = c->d_ss;
  H.s_opp = c->d_so;
  H.s_same32 = c->d_ss32;
  H.s_opp32 = c->d_so32;
  H.single_ok = c->d_sok;
  H.d1 = c->d_d1;
  H.d2 = c->d_d2;
  int na = (S.nelec + S.ms2) / 2, nb = (S.nelec - S.ms2) / 2;
  set_sector(H, na, nb);
  return MCDFCI_OK;
}

// Table sizing of src/run_wrapper.h:203-213: the largest power of two 2^p
// with 2^p * (8N + 16) bytes below max_memory GB; the reference's capacity
// is floor(0.79 * 2^p) and it stops when z_size > 0.79 * capacity.
void table_size(double max_memory, int nwords, uint64_t& slots, double& max_size, int& pw) {
  const int bytes = nwords * 8 + 16;
  uint64_t size = 8;
  while ((double)(size * (uint64_t)bytes) < max_memory * 1073741824.0) size *= 2;
  size /= 2;
  slots = size;
  max_size = (double)(uint64_t)(0.79 * (double)size);
  pw = 0;
  while ((1ULL << pw) < size) ++pw;
}

int alloc_iteration(mcdfci_ctx* c) {
  for (void* p : {(void*)c->pdet, (void*)c->pslot, (void*)c->px, (void*)c->pz, (void*)c->dx, (void*)c->Hs,
                  (void*)c->vis, (void*)c->stage, (void*)c->stage_cnt, (void*)c->part,
                  (void*)c->col_xz, (void*)c->cand, (void*)c->cand_cnt, (void*)c->pk0, (void*)c->pk1,
                  (void*)c->pc0, (void*)c->pc1, (void*)c->ctr, (void*)c->eslot, (void*)c->znorm})
    dfree(c, p);
  const int k = c->k, M = c->M;
  const DevHam& H = c->dh;
  c->nchunks = (H.n_enum + kChunk - 1) / kChunk;
  c->vis_words = c->nchunks * (kChunk / 64);
  const int mw = std::min(M, kChunk);
  if (c->nchunks > 1023) return fail(c, MCDFCI_E_UNSUPPORTED, "column too long for the merge kernel");
  if (M > kChunk) return fail(c, MCDFCI_E_UNSUPPORTED, "estimate above " + std::to_string(kChunk));
  int rc;
  if ((rc = dalloc(c, &c->pdet, k))) return rc;
  if ((rc = dalloc(c, &c->pslot, k))) return rc;
  if ((rc = dalloc(c, &c->px, k))) return rc;
  if ((rc = dalloc(c, &c->pz, k))) return rc;
  if ((rc = dalloc(c, &c->dx, k))) return rc;
  if ((rc = dalloc(c, &c->Hs, (size_t)k * k))) return rc;
  if ((rc = dalloc(c, &c->vis, (size_t)k * c->vis_words))) return rc;
  if ((rc = dalloc(c, &c->stage, (size_t)k * c->nchunks * mw))) return rc;
  if ((rc = dalloc(c, &c->stage_cnt, (size_t)k * c->nchunks))) return rc;
  if ((rc = dalloc(c, &c->part, (size_t)k * c->nchunks * 4))) return rc;
  if ((rc = dalloc(c, &c->col_xz, (size_t)2 * k))) return rc;
  if ((rc = dalloc(c, &c->cand, (size_t)k * M))) return rc;
  if ((rc = dalloc(c, &c->cand_cnt, k))) return rc;
  if ((rc = dalloc(c, &c->pk0, (size_t)k * k))) return rc;  // pick-tree lists of up to k entries
  if ((rc = dalloc(c, &c->pk1, (size_t)k * k))) return rc;
  if ((rc = dalloc(c, &c->pc0, k))) return rc;
  if ((rc = dalloc(c, &c->pc1, k))) return rc;
  if ((rc = dalloc(c, &c->ctr, (size_t)kCtrStripes * kCtrStride))) return rc;
  if ((rc = dalloc(c, &c->eslot, (size_t)k * c->vis_words * 64))) return rc;
  if ((rc = dalloc(c, &c->znorm, (size_t)kZNormCap))) return rc;
  HIPCHECK(c, hipMemsetAsync(c->cand_cnt, 0, sizeof(int) * k, c->stream));
  HIPCHECK(c, hipMemsetAsync(c->ctr, 0, sizeof(long long) * kCtrStripes * kCtrStride, c->stream));
  return MCDFCI_OK;
}

int check_state_error(mcdfci_ctx* c, const DevState& s) {
  switch (s.error) {
    case ERR_NONE:
      return MCDFCI_OK;
    case ERR_TABLE_FULL:
      return fail(c, MCDFCI_E_TABLE_FULL,
                  "The hash table is full. Please increase max_wavefunction_size or z_threshold.");
    case ERR_NUMERIC:
      return fail(c, MCDFCI_E_NUMERIC,
                  "The lowest subspace eigenvalue is not negative, so sqrt(-lambda) is undefined. The "
                  "Hamiltonian (core energy included) must have a negative lowest eigenvalue.");
    case ERR_CANDIDATES:
      return fail(c, MCDFCI_E_CANDIDATES,
                  "Fewer than num_of_coordinate distinct candidate determinants are available.");
    case ERR_PROBE:
      return fail(c, MCDFCI_E_TABLE_FULL, "The hash table has no free slot left.");
    case ERR_CAP:
      return fail(c, MCDFCI_E_UNSUPPORTED, "candidate list exceeds the device sort capacity.");
    case ERR_SYNC:
      return fail(c, MCDFCI_E_HIP, "the subspace solver's workgroups lost an exchanged value (no reply in time).");
    default:
      return fail(c, MCDFCI_E_HIP, "device error " + std::to_string(s.error));
  }
}

// The x field of slot 0 and the slot stride in doubles: k_eig writes the
// new x of each picked coordinate through its slot index.
double* tab_x(mcdfci_ctx* c) {
  return c->nw == 1 ? &((SlotW<1>*)c->tab)->x : &((SlotW<2>*)c->tab)->x;
}
int slot_doubles(int nw) { return (int)(slot_bytes(nw) / sizeof(double)); }

void launch_update_z(mcdfci_ctx* c, int k, int count_iteration) {
  hipStream_t s = c->stream;
  launch_scatter(s, c->nw, c->dh, c->st, c->pdet, c->dx, c->Hs, c->k, c->tab, c->nslots - 1, c->vis, c->vis_words,
                 k, c->scatter_ept, c->ctr, c->eslot);
  if (c->timing && count_iteration) (void)hipEventRecord(c->ev[4], s);
  launch_pass2(s, c->nw, c->dh, c->st, c->pdet, c->dx, c->Hs, c->k, c->tab, c->nslots - 1, c->vis, c->vis_words,
               c->M, k, c->nchunks, c->stage, c->stage_cnt, c->part, c->eslot, c->znorm);
  if (c->timing && count_iteration) (void)hipEventRecord(c->ev[5], s);
  launch_merge(s, c->nw, c->st, c->pslot, c->tab, c->M, k, c->nchunks, c->stage, c->stage_cnt, c->part,
               c->col_xz, c->cand, c->cand_cnt, c->dx, count_iteration, c->ctr, c->nslots, c->znorm,
               c->merge_tails);
}

// One iteration of src/solver.h:701-759.  With timing enabled every kernel
// is bracketed by HIP events on the solver stream and the iteration is
// synchronised to read them.
void launch_iteration(mcdfci_ctx* c) {
  hipStream_t s = c->stream;
  const int k = c->k;
  if (c->timing) (void)hipEventRecord(c->ev[0], s);
  launch_pick(s, c->nw, c->st, k, c->M, c->cand, c->cand_cnt, c->tab, c->pdet, c->pslot, c->px, c->pz, c->pk0,
              c->pk1, c->pc0, c->pc1, c->pick_tree);
  if (c->timing) (void)hipEventRecord(c->ev[1], s);
  launch_submatrix(s, c->nw, c->dh, c->st, c->pdet, k, c->Hs, k);
  if (c->timing) (void)hipEventRecord(c->ev[2], s);
  if (c->opts.coordinate_update == MCDFCI_UPDATE_LS)
    launch_ls(s, c->st, c->Hs, c->px, c->pz, c->dx, tab_x(c), slot_doubles(c->nw), c->pslot);
  else
    launch_eig(s, c->st, k, k, c->Hs, c->px, c->pz, c->dx, tab_x(c), slot_doubles(c->nw), c->pslot, 0, c->xbuf);
  if (c->timing) (void)hipEventRecord(c->ev[3], s);
  launch_update_z(c, k, 1);
  if (c->timing) {
    (void)hipEventRecord(c->ev[6], s);
    (void)hipEventSynchronize(c->ev[6]);
    float ms[6];
    for (int i = 0; i < 6; ++i) (void)hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]);
    for (int i = 0; i < 6; ++i) c->ktimes[i] += ms[i] * 1e-3;
    c->ktimed += 1;
    // step_timer laps: pick, submatrix, columns (fused into update_z), eig
    // (with update_x), update_x, update_z
    c->timers[0] += ms[0] * 1e-3;
    c->timers[1] += ms[1] * 1e-3;
    c->timers[3] += ms[2] * 1e-3;
    c->timers[5] += (ms[3] + ms[4] + ms[5]) * 1e-3;
  }
}

void drop_graph(mcdfci_ctx* c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
}

// Captures graph_iters iterations of the inner loop on the solver stream.
int capture_graph(mcdfci_ctx* c) {
  drop_graph(c);
  HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < c->graph_iters; ++i) launch_iteration(c);
  hipGraph_t g = nullptr;
  HIPCHECK(c, hipStreamEndCapture(c->stream, &g));
  hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->graph = nullptr;
    return fail(c, MCDFCI_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return MCDFCI_OK;
}

double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  return v.empty() ? 0.0 : v[v.size() / 2];
}

}  // namespace

extern "C" {

const char* mcdfci_last_error(const mcdfci_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

}  // extern "C"

namespace {

// Context creation from integrals already in SpatialHam form (restricted or
// unrestricted); every public create entry point ends here.
int create_ctx(SpatialHam&& ham_in, const mcdfci_opts* opts, mcdfci_ctx** out) {
  const SpatialHam* ham = &ham_in;
  if (ham->norb > kMaxOrb2)
    return fail(nullptr, MCDFCI_E_UNSUPPORTED,
                "the device path supports up to " + std::to_string(kMaxOrb2) + " spatial orbitals");
  if (opts->num_of_coordinate < 1 || opts->num_of_coordinate > kMaxK)
    return fail(nullptr, MCDFCI_E_UNSUPPORTED, "num_of_coordinate must be in [1, " + std::to_string(kMaxK) + "]");
  if (opts->estimate < 1) return fail(nullptr, MCDFCI_E_INPUT, "estimate must be positive");
  if (!(opts->max_memory > 0.0))
    return fail(nullptr, MCDFCI_E_INPUT,
                "solver:cdfci:max_memory is invalid. It is the maximum memory (GB) allowed for the "
                "wavefunction. Please set a positive number.");
  if (opts->threshold < 0.0) return fail(nullptr, MCDFCI_E_UNSUPPORTED, "hamiltonian:threshold must be >= 0");
  if (opts->coordinate_update != MCDFCI_UPDATE_EIG && opts->coordinate_update != MCDFCI_UPDATE_LS)
    return fail(nullptr, MCDFCI_E_INPUT, "coordinate_update must be MCDFCI_UPDATE_EIG or MCDFCI_UPDATE_LS");
  if (opts->coordinate_update == MCDFCI_UPDATE_LS && opts->num_of_coordinate != 1)
    return fail(nullptr, MCDFCI_E_UNSUPPORTED,
                "the line-search coordinate update takes num_of_coordinate = 1 (one coordinate per iteration)");
  {
    // Slot indices are 32-bit (staging lists, candidates, probe ends), so
    // the table holds at most 2^32 slots; refuse before allocating.
    uint64_t slots;
    double msz;
    int pw;
    table_size(opts->max_memory, 1 + (2 * ham->norb - 1) / 64, slots, msz, pw);
    if (slots > (1ULL << 32))
      return fail(nullptr, MCDFCI_E_UNSUPPORTED,
                  "max_memory gives " + std::to_string(slots) +
                      " wavefunction slots; the device table holds at most 2^32 (lower max_memory)");
  }
  mcdfci_ctx* c = new mcdfci_ctx();
  c->opts = *opts;
  if (const char* zb = std::getenv("MCDFCI_ZNORM_BITS")) {
    // test hook: fold z word B into A from 2^bits on (53..61) instead of 2^61
    const int v = std::atoi(zb);
    if (v >= 53 && v <= 61) c->znorm_bits = v;
  }
  if (const char* mt = std::getenv("MCDFCI_MERGE_TAILS")) {
    // test hook: the column merge falls back to the full tournament when more
    // than this many (0..7) chunk lists can reach the column's top M
    const int v = std::atoi(mt);
    if (v >= 0 && v <= 7) c->merge_tails = v;
  }
  c->k = opts->num_of_coordinate;
  c->M = opts->estimate;
  c->device = opts->device;
  c->nw = ham->norb <= kMaxOrb ? 1 : 2;
  c->nio = 1 + (2 * ham->norb - 1) / 64;
  c->ham = std::move(ham_in);
  int rc = MCDFCI_OK;
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) rc = fail(c, MCDFCI_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  if (!rc) {
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(c, MCDFCI_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  if (!rc) rc = build_tables(c);
  if (!rc) {
    // Slot count and capacity follow the reference's sizing for its word
    // count; each slot here is 32 (one word) or 40 (two words) bytes.
    table_size(opts->max_memory, c->nio, c->nslots, c->max_size, c->table_pow);
    uint8_t* t = nullptr;
    rc = dalloc(c, &t, c->nslots * slot_bytes(c->nw));
    c->tab = t;
  }
  if (!rc) rc = dalloc(c, &c->st, 1);
  if (!rc) rc = dalloc(c, &c->st_scratch, 1);
  if (!rc) rc = dalloc(c, &c->xbuf, kXchgWords);
  if (!rc && hipMemset(c->xbuf, 0, sizeof(uint64_t) * kXchgWords) != hipSuccess)
    rc = fail(c, MCDFCI_E_HIP, "hipMemset of the exchange buffer failed");
  if (!rc) {
    // Address of this device's `scratch`, looked up once per device (the
    // device is current here).
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(c->device);
    if (it == g_scratch_addr.end()) {
      void* p = nullptr;
      hipError_t e2 = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e2 != hipSuccess || !p)
        rc = fail(c, MCDFCI_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e2));
      else
        it = g_scratch_addr.emplace(c->device, (double*)p).first;
    }
    if (!rc) c->noalloc = it->second;
  }
  if (!rc)
    for (int i = 0; i < 7 && !rc; ++i)
      if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = fail(c, MCDFCI_E_HIP, "hipEventCreate failed");
  if (rc) {
    g_err = c->err;
    mcdfci_destroy(c);
    return rc;
  }
  *out = c;
  return MCDFCI_OK;
}

}  // namespace

extern "C" {

int mcdfci_create(const mcdfci_ham* ham, const mcdfci_opts* opts, mcdfci_ctx** out) {
  if (!ham || !opts || !out) return fail(nullptr, MCDFCI_E_INPUT, "null argument");
  *out = nullptr;
  if (ham->norb <= 0 || !ham->h || !ham->eri) return fail(nullptr, MCDFCI_E_INPUT, "invalid Hamiltonian");
  if (ham->norb > kMaxOrb2)
    return fail(nullptr, MCDFCI_E_UNSUPPORTED,
                "the device path supports up to " + std::to_string(kMaxOrb2) + " spatial orbitals");
  SpatialHam S;
  S.norb = ham->norb;
  S.nelec = ham->nelec;
  S.ms2 = ham->ms2;
  S.core = ham->core;
  const size_t n2 = (size_t)S.norb * S.norb;
  S.h.assign(ham->h, ham->h + n2);
  S.eri.assign(ham->eri, ham->eri + n2 * n2);
  if (ham->orbital_energy)
    for (int p = 0; p < S.norb; ++p) {
      S.orb_e.push_back({2 * p, ham->orbital_energy[p]});
      S.orb_e.push_back({2 * p + 1, ham->orbital_energy[p]});
    }
  return create_ctx(std::move(S), opts, out);
}

int mcdfci_create_uhf(const mcdfci_ham_uhf* ham, const mcdfci_opts* opts, mcdfci_ctx** out) {
  if (!ham || !opts || !out) return fail(nullptr, MCDFCI_E_INPUT, "null argument");
  *out = nullptr;
  if (ham->norb <= 0 || !ham->h_alpha || !ham->h_beta || !ham->eri_aa || !ham->eri_bb || !ham->eri_ab)
    return fail(nullptr, MCDFCI_E_INPUT, "invalid Hamiltonian");
  if (ham->norb > kMaxOrb2)
    return fail(nullptr, MCDFCI_E_UNSUPPORTED,
                "the device path supports up to " + std::to_string(kMaxOrb2) + " spatial orbitals");
  SpatialHam S;
  S.norb = ham->norb;
  S.nelec = ham->nelec;
  S.ms2 = ham->ms2;
  S.core = ham->core;
  S.uhf = true;
  const size_t n2 = (size_t)S.norb * S.norb, n4 = n2 * n2;
  S.h.assign(ham->h_alpha, ham->h_alpha + n2);
  S.hb.assign(ham->h_beta, ham->h_beta + n2);
  S.eri.assign(ham->eri_aa, ham->eri_aa + n4);
  S.eri_bb.assign(ham->eri_bb, ham->eri_bb + n4);
  S.eri_ab.assign(ham->eri_ab, ham->eri_ab + n4);
  if (ham->orbital_energy)
    for (int i = 0; i < 2 * S.norb; ++i) S.orb_e.push_back({i, ham->orbital_energy[i]});
  return create_ctx(std::move(S), opts, out);
}

int mcdfci_create_from_fcidump(const char* path, const mcdfci_opts* opts, mcdfci_ctx** out) {
  if (!path || !opts || !out) return fail(nullptr, MCDFCI_E_INPUT, "null argument");
  *out = nullptr;
  SpatialHam S;
  try {
    S = read_fcidump(path, nullptr);
  } catch (const std::exception& ex) {
    return fail(nullptr, MCDFCI_E_INPUT, ex.what());
  }
  return create_ctx(std::move(S), opts, out);
}

int mcdfci_destroy(mcdfci_ctx* c) {
  if (!c) return MCDFCI_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return MCDFCI_OK;
}

int mcdfci_info(const mcdfci_ctx* c, int64_t* info) {
  if (!c || !info) return fail(nullptr, MCDFCI_E_INPUT, "null argument");
  info[0] = c->nio;
  info[1] = 2 * c->ham.norb;
  info[2] = c->ham.nelec;
  info[3] = c->ham.ms2;
  info[4] = c->dh.n_enum;
  info[5] = c->k;
  info[6] = c->ham.orb_e.empty() ? 0 : 1;
  info[7] = c->ham.uhf ? 1 : 0;
  return MCDFCI_OK;
}

int mcdfci_hartree_fock(mcdfci_ctx* c, uint64_t* det) {
  if (!c || !det) return fail(c, MCDFCI_E_INPUT, "null argument");
  uint64_t d[2];
  hartree_fock(c->ham, d, nullptr);
  store_det(det, D2{d[0], d[1]}, c->nio);
  return MCDFCI_OK;
}

int mcdfci_init(mcdfci_ctx* c, const uint64_t* ref_det) {
  if (!c) return fail(nullptr, MCDFCI_E_INPUT, "null context");
  HIPCHECK(c, hipSetDevice(c->device));
  D2 det;
  if (ref_det) {
    det = load_det(ref_det, c->nio);
  } else {
    uint64_t d[2];
    hartree_fock(c->ham, d, nullptr);
    det = D2{d[0], d[1]};
  }
  int rc = check_det(c, det, "reference determinant");
  if (rc) return rc;
  int na, nb;
  det_sector(det, na, nb);
  if (na + nb != c->ham.nelec)
    return fail(c, MCDFCI_E_INPUT,
                "Reference determinant is invalid. The number of valid orbitals does not match the number of "
                "electrons.");
  set_sector(c->dh, na, nb);
  drop_graph(c);  // the iteration buffers are reallocated
  rc = alloc_iteration(c);
  if (rc) return rc;
  launch_table_init(c->stream, c->nw, c->tab, c->nslots);
  DevState s{};
  s.z_thr = c->opts.z_threshold;
  s.max_size = c->max_size;
  s.znorm_bits = c->znorm_bits;
  HIPCHECK(c, hipMemcpyAsync(c->st, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  launch_init_ref(c->stream, c->nw, c->st, det, c->tab, c->nslots - 1, c->pdet, c->pslot, c->px, c->pz, c->dx);
  launch_submatrix(c->stream, c->nw, c->dh, c->st, c->pdet, 1, c->Hs, c->k);
  launch_update_z(c, 1, 0);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  for (double& t : c->timers) t = 0.0;
  for (double& t : c->ktimes) t = 0.0;
  c->ktimed = 0;
  c->ready = true;
  return check_state_error(c, s);
}

int mcdfci_set_z_threshold(mcdfci_ctx* c, double z) {
  if (!c) return fail(nullptr, MCDFCI_E_INPUT, "null context");
  c->opts.z_threshold = z;
  return MCDFCI_OK;
}

int mcdfci_step(mcdfci_ctx* c, int64_t n) {
  if (!c) return fail(nullptr, MCDFCI_E_INPUT, "null context");
  if (!c->ready) return fail(c, MCDFCI_E_INPUT, "mcdfci_init has not been called");
  HIPCHECK(c, hipSetDevice(c->device));
  if (c->graph_iters > 0 && !c->timing && n >= c->graph_iters) {
    if (!c->graph) {
      int rc = capture_graph(c);
      if (rc) return rc;
    }
    for (; n >= c->graph_iters; n -= c->graph_iters) HIPCHECK(c, hipGraphLaunch(c->graph, c->stream));
  }
  for (int64_t i = 0; i < n; ++i) launch_iteration(c);
  HIPCHECK(c, hipGetLastError());
  return MCDFCI_OK;
}

int mcdfci_access_ceiling(mcdfci_ctx* c, int mode, int64_t n_ops, double* seconds) {
  if (!c || !seconds || mode < 0 || mode > 4 || n_ops < 1) return fail(c, MCDFCI_E_INPUT, "invalid argument");
  HIPCHECK(c, hipSetDevice(c->device));
  unsigned long long* sink = nullptr;
  int rc = dalloc(c, &sink, 1);
  if (rc) return rc;
  c->ready = false;  // the table no longer holds a wavefunction
  drop_graph(c);
  // one untimed pass brings the clocks up; the second is timed
  for (int pass = 0; pass < 2; ++pass) {
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    launch_access_probe(c->stream, c->nw, c->tab, c->nslots - 1, (long long)n_ops, mode,
                        0x9e3779b97f4a7c15ULL * (uint64_t)(pass + 1), sink);
    HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
  }
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipEventSynchronize(c->ev[1]));
  float ms = 0.0f;
  HIPCHECK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  dfree(c, sink);
  *seconds = ms * 1e-3;
  return MCDFCI_OK;
}

}  // extern "C"

namespace {

// Checkpoint file: a fixed header, the raw device state, the candidate
// lists, then the table's entries in chunks of at most kCkChunk (per chunk:
// the count, two words per determinant, x, and the (zA, zB) pairs).
constexpr char kCkMagic[8] = {'M', 'C', 'D', 'F', 'C', 'I', 'C', 'K'};
constexpr int kCkVersion = 3;
constexpr long long kCkChunk = 1LL << 22;

struct CkHeader {
  char magic[8];
  int32_t version, state_bytes, nio, nw;
  int32_t norb, nelec, ms2, na, nb, k, M, precision, coordinate_update, pad;
  double threshold, core;
  uint64_t ham_hash;
  int64_t entries;
};

uint64_t ham_hash(const SpatialHam& S) {
  uint64_t h = 0xcbf29ce484222325ULL;
  auto feed = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001b3ULL;
  };
  feed(S.h.data(), S.h.size() * sizeof(double));
  feed(S.eri.data(), S.eri.size() * sizeof(double));
  feed(&S.core, sizeof(S.core));
  if (S.uhf) {
    feed(S.hb.data(), S.hb.size() * sizeof(double));
    feed(S.eri_bb.data(), S.eri_bb.size() * sizeof(double));
    feed(S.eri_ab.data(), S.eri_ab.size() * sizeof(double));
  }
  return h;
}

struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

}  // namespace

extern "C" {

int mcdfci_checkpoint_save(mcdfci_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MCDFCI_E_INPUT, "null argument");
  if (!c->ready) return fail(c, MCDFCI_E_INPUT, "mcdfci_init has not been called");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  DevState s;
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  if (s.error) return check_state_error(c, s);
  // count, then compact the occupied slots
  unsigned long long* cnt = nullptr;
  int rc = dalloc(c, &cnt, 1);
  if (rc) return rc;
  HIPCHECK(c, hipMemsetAsync(cnt, 0, 8, c->stream));
  launch_table_compact(c->stream, c->nw, c->tab, c->nslots, nullptr, nullptr, nullptr, cnt, 0);
  unsigned long long m = 0;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(&m, cnt, 8, hipMemcpyDeviceToHost));
  uint64_t* od = nullptr;
  double* ox = nullptr;
  int64_t* oz = nullptr;
  auto release = [&]() {
    for (void* p : {(void*)cnt, (void*)od, (void*)ox, (void*)oz}) dfree(c, p);
  };
  if ((rc = dalloc(c, &od, 2 * m)) || (rc = dalloc(c, &ox, m)) || (rc = dalloc(c, &oz, 2 * m))) {
    release();
    return rc;
  }
  HIPCHECK(c, hipMemsetAsync(cnt, 0, 8, c->stream));
  launch_table_compact(c->stream, c->nw, c->tab, c->nslots, od, ox, oz, cnt, (long long)m);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  std::vector<int> ccnt(c->k);
  std::vector<CandEnt> cand((size_t)c->k * c->M);
  HIPCHECK(c, hipMemcpy(ccnt.data(), c->cand_cnt, sizeof(int) * c->k, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(cand.data(), c->cand, sizeof(CandEnt) * cand.size(), hipMemcpyDeviceToHost));
  FileCloser fc{fopen(path, "wb")};
  if (!fc.f) {
    release();
    return fail(c, MCDFCI_E_INPUT, std::string("cannot open ") + path + " for writing");
  }
  CkHeader hd{};
  std::memcpy(hd.magic, kCkMagic, 8);
  hd.version = kCkVersion;
  hd.state_bytes = (int32_t)sizeof(DevState);
  hd.nio = c->nio;
  hd.nw = c->nw;
  hd.norb = c->ham.norb;
  hd.nelec = c->ham.nelec;
  hd.ms2 = c->ham.ms2;
  hd.na = c->dh.na;
  hd.nb = c->dh.nb;
  hd.k = c->k;
  hd.M = c->M;
  hd.precision = c->opts.precision;
  hd.coordinate_update = c->opts.coordinate_update;
  hd.threshold = c->opts.threshold;
  hd.core = c->ham.core;
  hd.ham_hash = ham_hash(c->ham);
  hd.entries = (int64_t)m;
  bool ok = fwrite(&hd, sizeof(hd), 1, fc.f) == 1 && fwrite(&s, sizeof(s), 1, fc.f) == 1 &&
            fwrite(ccnt.data(), sizeof(int), ccnt.size(), fc.f) == ccnt.size() &&
            fwrite(cand.data(), sizeof(CandEnt), cand.size(), fc.f) == cand.size();
  std::vector<uint64_t> hd2;
  std::vector<double> hx;
  std::vector<int64_t> hz;
  for (long long off = 0; ok && off < (long long)m; off += kCkChunk) {
    const long long n = std::min<long long>(kCkChunk, (long long)m - off);
    hd2.resize(2 * n);
    hx.resize(n);
    hz.resize(2 * n);
    if (hipMemcpy(hd2.data(), od + 2 * off, 16 * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hx.data(), ox + off, 8 * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hz.data(), oz + 2 * off, 16 * n, hipMemcpyDeviceToHost) != hipSuccess) {
      release();
      return fail(c, MCDFCI_E_HIP, "checkpoint copy failed");
    }
    const int64_t nn = n;
    ok = fwrite(&nn, sizeof(nn), 1, fc.f) == 1 && fwrite(hd2.data(), 8, 2 * n, fc.f) == (size_t)(2 * n) &&
         fwrite(hx.data(), 8, n, fc.f) == (size_t)n && fwrite(hz.data(), 8, 2 * n, fc.f) == (size_t)(2 * n);
  }
  release();
  if (!ok) return fail(c, MCDFCI_E_INPUT, std::string("write to ") + path + " failed");
  return MCDFCI_OK;
}

int mcdfci_checkpoint_load(mcdfci_ctx* c, const char* path) {
  if (!c || !path) return fail(c, MCDFCI_E_INPUT, "null argument");
  HIPCHECK(c, hipSetDevice(c->device));
  FileCloser fc{fopen(path, "rb")};
  if (!fc.f) return fail(c, MCDFCI_E_INPUT, std::string("cannot open ") + path);
  CkHeader hd{};
  DevState s{};
  if (fread(&hd, sizeof(hd), 1, fc.f) != 1 || std::memcmp(hd.magic, kCkMagic, 8) != 0 ||
      hd.version != kCkVersion || hd.state_bytes != (int32_t)sizeof(DevState))
    return fail(c, MCDFCI_E_INPUT, std::string(path) + " is not a checkpoint of this library version");
  if (hd.nio != c->nio || hd.nw != c->nw || hd.norb != c->ham.norb || hd.nelec != c->ham.nelec ||
      hd.ms2 != c->ham.ms2 || hd.k != c->k || hd.M != c->M || hd.precision != c->opts.precision ||
      hd.coordinate_update != c->opts.coordinate_update || hd.threshold != c->opts.threshold ||
      hd.ham_hash != ham_hash(c->ham))
    return fail(c, MCDFCI_E_INPUT,
                "the checkpoint was written for another Hamiltonian or other num_of_coordinate / estimate / "
                "precision / coordinate_update / threshold options");
  if (hd.na < 0 || hd.nb < 0 || hd.na + hd.nb != c->ham.nelec || hd.na > c->ham.norb || hd.nb > c->ham.norb)
    return fail(c, MCDFCI_E_INPUT, "the checkpoint's electron sector does not match NELEC");
  if (fread(&s, sizeof(s), 1, fc.f) != 1) return fail(c, MCDFCI_E_INPUT, "truncated checkpoint");
  std::vector<int> ccnt(c->k);
  std::vector<CandEnt> cand((size_t)c->k * c->M);
  if (fread(ccnt.data(), sizeof(int), ccnt.size(), fc.f) != ccnt.size() ||
      fread(cand.data(), sizeof(CandEnt), cand.size(), fc.f) != cand.size())
    return fail(c, MCDFCI_E_INPUT, "truncated checkpoint");
  if ((double)s.z_size > 0.79 * c->max_size || hd.entries > (int64_t)(0.79 * c->max_size) + c->k + 16)
    return fail(c, MCDFCI_E_TABLE_FULL, "the checkpoint does not fit this context's wavefunction table");
  c->ready = false;
  set_sector(c->dh, hd.na, hd.nb);
  drop_graph(c);
  int rc = alloc_iteration(c);
  if (rc) return rc;
  launch_table_init(c->stream, c->nw, c->tab, c->nslots);
  uint64_t* dd2 = nullptr;
  double* dxx = nullptr;
  int64_t* dz2 = nullptr;
  int* dfail = nullptr;
  auto release = [&]() {
    for (void* p : {(void*)dd2, (void*)dxx, (void*)dz2, (void*)dfail}) dfree(c, p);
  };
  const long long cap = std::min<long long>(kCkChunk, std::max<long long>(hd.entries, 1));
  if ((rc = dalloc(c, &dd2, 2 * cap)) || (rc = dalloc(c, &dxx, cap)) || (rc = dalloc(c, &dz2, 2 * cap)) ||
      (rc = dalloc(c, &dfail, 1))) {
    release();
    return rc;
  }
  HIPCHECK(c, hipMemsetAsync(dfail, 0, sizeof(int), c->stream));
  std::vector<uint64_t> hd2;
  std::vector<double> hx;
  std::vector<int64_t> hz;
  for (long long off = 0; off < hd.entries;) {
    int64_t n = 0;
    if (fread(&n, sizeof(n), 1, fc.f) != 1 || n < 1 || n > cap || off + n > hd.entries) {
      release();
      return fail(c, MCDFCI_E_INPUT, "truncated or corrupt checkpoint");
    }
    hd2.resize(2 * n);
    hx.resize(n);
    hz.resize(2 * n);
    if (fread(hd2.data(), 8, 2 * n, fc.f) != (size_t)(2 * n) || fread(hx.data(), 8, n, fc.f) != (size_t)n ||
        fread(hz.data(), 8, 2 * n, fc.f) != (size_t)(2 * n)) {
      release();
      return fail(c, MCDFCI_E_INPUT, "truncated checkpoint");
    }
    // the copies are synchronous, so the host buffers may be refilled next
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(dd2, hd2.data(), 16 * n, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(dxx, hx.data(), 8 * n, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(dz2, hz.data(), 16 * n, hipMemcpyHostToDevice));
    launch_table_insert(c->stream, c->nw, c->tab, c->nslots - 1, n, dd2, dxx, dz2, dfail);
    off += n;
  }
  HIPCHECK(c, hipMemcpyAsync(c->cand_cnt, ccnt.data(), sizeof(int) * c->k, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(c->cand, cand.data(), sizeof(CandEnt) * cand.size(), hipMemcpyHostToDevice,
                             c->stream));
  launch_cand_resolve(c->stream, c->nw, c->tab, c->nslots - 1, c->cand, c->cand_cnt, c->k, c->M, dfail);
  s.max_size = c->max_size;
  s.ticket = 0;
  HIPCHECK(c, hipMemcpyAsync(c->st, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  int f = 0;
  HIPCHECK(c, hipMemcpy(&f, dfail, sizeof(int), hipMemcpyDeviceToHost));
  release();
  if (f) return fail(c, MCDFCI_E_TABLE_FULL, "the checkpoint's entries do not fit the table");
  c->opts.z_threshold = s.z_thr;
  for (double& t : c->timers) t = 0.0;
  for (double& t : c->ktimes) t = 0.0;
  c->ktimed = 0;
  c->ready = true;
  return MCDFCI_OK;
}

int mcdfci_set_graph(mcdfci_ctx* c, int iterations_per_graph) {
  if (!c || iterations_per_graph < 0) return fail(c, MCDFCI_E_INPUT, "invalid argument");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  if (iterations_per_graph != c->graph_iters) drop_graph(c);
  c->graph_iters = iterations_per_graph;
  return MCDFCI_OK;
}

// Online tuning on real iterations: each round runs one iteration per
// (elements per thread, pick form) pair, timing the scatter and the pick
// with HIP events; the per-kernel medians pick the configuration.
int mcdfci_autotune(mcdfci_ctx* c, int rounds, int64_t* choice) {
  if (!c || rounds < 1) return fail(c, MCDFCI_E_INPUT, "invalid argument");
  if (!c->ready) return fail(c, MCDFCI_E_INPUT, "mcdfci_init has not been called");
  HIPCHECK(c, hipSetDevice(c->device));
  constexpr int kNE = 5;
  const int epts[kNE] = {1, 2, 4, 8, 16};
  const bool can_single = (long long)c->k * c->M <= kSortCap;
  const int npick = can_single ? 2 : 1;
  std::vector<double> ts[kNE], tp[2];
  const bool timing = c->timing;
  double kt[6];
  std::copy(c->ktimes, c->ktimes + 6, kt);
  const long long kn = c->ktimed;
  c->timing = true;
  int64_t iters = 0;
  for (int r = 0; r < rounds; ++r)
    for (int e = 0; e < kNE; ++e)
      for (int p = 0; p < npick; ++p) {
        c->scatter_ept = epts[e];
        c->pick_tree = !can_single || p == 1;
        const double s0 = c->ktimes[3], p0 = c->ktimes[0];
        launch_iteration(c);  // synchronises (timing mode)
        ts[e].push_back(c->ktimes[3] - s0);
        tp[c->pick_tree ? 1 : 0].push_back(c->ktimes[0] - p0);
        ++iters;
      }
  c->timing = timing;
  // the tuning iterations count in the step timers but not in the
  // per-kernel times of the chosen configuration
  std::copy(kt, kt + 6, c->ktimes);
  c->ktimed = kn;
  int best = 2;
  for (int e = 0; e < kNE; ++e)
    if (median(ts[e]) < median(ts[best])) best = e;
  c->scatter_ept = epts[best];
  c->pick_tree = !can_single || median(tp[1]) < median(tp[0]);
  drop_graph(c);
  HIPCHECK(c, hipGetLastError());
  if (choice) {
    choice[0] = c->scatter_ept;
    choice[1] = c->pick_tree ? 1 : 0;
    choice[2] = iters;
  }
  DevState s;
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  return check_state_error(c, s);
}

int mcdfci_set_launch(mcdfci_ctx* c, int elements_per_thread, int tree_pick) {
  if (!c) return MCDFCI_E_INPUT;
  const int e = elements_per_thread;
  if (!(e == 1 || e == 2 || e == 4 || e == 8 || e == 16) || (tree_pick != 0 && tree_pick != 1))
    return fail(c, MCDFCI_E_INPUT, "invalid launch form");
  if (!tree_pick && (long long)c->k * c->M > kSortCap)
    return fail(c, MCDFCI_E_INPUT, "num_of_coordinate * estimate needs the merge-tree pick");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  c->scatter_ept = e;
  c->pick_tree = tree_pick != 0;
  drop_graph(c);
  return MCDFCI_OK;
}

int mcdfci_report_get(mcdfci_ctx* c, mcdfci_report* r) {
  if (!c || !r) return fail(c, MCDFCI_E_INPUT, "null argument");
  if (!c->ready) {
    // Before mcdfci_init only the sizing fields are meaningful.
    std::memset(r, 0, sizeof(*r));
    r->table_slots = (int64_t)c->nslots;
    r->max_size = c->max_size;
    return MCDFCI_OK;
  }
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  DevState s;
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  i128 xz = (((i128)s.xz_hi) << 64) | (i128)s.xz_lo;
  dd xx{s.xx_hi, s.xx_lo};
  std::memset(r, 0, sizeof(*r));
  r->energy = dd_div(fix_to_dd(xz), xx).hi;
  r->dx_max = s.dx_max;
  r->x_size = s.x_size;
  r->z_size = s.z_size;
  r->cols_size = s.cols_size;
  r->updates = s.updates;
  r->scale = s.sc_hi;
  r->xx = s.xx_hi;
  r->xz = fix_to_d(xz);
  r->iterations = s.iterations;
  r->max_residual = s.max_residual;
  r->curr_max_abs_evec = s.curr_max_abs_evec;
  r->sum_iterations = s.sum_iterations;
  r->sum_inv_iter = s.sum_inv_iter;
  r->curr_inv_iter = s.curr_inv_iter;
  r->sum_diff_inv_iter = s.sum_diff_inv;
  r->curr_diff_inv_iter = s.curr_diff_inv;
  r->max_diff_inv_iter = s.max_diff_inv;
  r->table_slots = (int64_t)c->nslots;
  r->max_size = c->max_size;
  r->eig_unconverged = s.eig_unconverged;
  {
    std::vector<int> cnt(c->k);
    HIPCHECK(c, hipMemcpy(cnt.data(), c->cand_cnt, sizeof(int) * c->k, hipMemcpyDeviceToHost));
    int64_t tot = 0;
    for (int t = 0; t < s.k_prev && t < c->k; ++t) tot += cnt[t];
    r->candidates = tot;
  }
  return check_state_error(c, s);
}

int mcdfci_set_timing(mcdfci_ctx* c, int enable) {
  if (!c) return fail(nullptr, MCDFCI_E_INPUT, "null context");
  c->timing = enable != 0;
  return MCDFCI_OK;
}

int mcdfci_timers(mcdfci_ctx* c, double* t6) {
  if (!c || !t6) return fail(c, MCDFCI_E_INPUT, "null argument");
  for (int i = 0; i < 6; ++i) t6[i] = c->timers[i];
  return MCDFCI_OK;
}

int mcdfci_kernel_times(mcdfci_ctx* c, double* t6, int64_t* iterations) {
  if (!c || !t6) return fail(c, MCDFCI_E_INPUT, "null argument");
  for (int i = 0; i < 6; ++i) t6[i] = c->ktimes[i];
  if (iterations) *iterations = c->ktimed;
  return MCDFCI_OK;
}

int mcdfci_column(mcdfci_ctx* c, const uint64_t* det, uint64_t* dets, double* vals, int64_t cap, int64_t* n) {
  if (!c || !det || !n) return fail(c, MCDFCI_E_INPUT, "null argument");
  HIPCHECK(c, hipSetDevice(c->device));
  DevHam H = c->dh;
  const D2 d0 = load_det(det, c->nio);
  int rc = check_det(c, d0, "determinant");
  if (rc) return rc;
  int na, nb;
  det_sector(d0, na, nb);
  set_sector(H, na, nb);
  D2* dd_ = nullptr;
  double* dh_ = nullptr;
  double* dv = nullptr;
  unsigned long long* cnt = nullptr;
  if ((rc = dalloc(c, &dd_, 1))) return rc;
  if ((rc = dalloc(c, &dh_, 1))) return rc;
  const long long ncap = H.n_enum;
  uint64_t* od = nullptr;
  if ((rc = dalloc(c, &od, 2 * ncap))) return rc;
  if ((rc = dalloc(c, &dv, ncap))) return rc;
  if ((rc = dalloc(c, &cnt, 1))) return rc;
  HIPCHECK(c, hipMemcpy(dd_, &d0, sizeof(D2), hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemsetAsync(cnt, 0, 8, c->stream));
  launch_submatrix(c->stream, c->nw, H, nullptr, dd_, 1, dh_, 1);
  double diag;
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(&diag, dh_, 8, hipMemcpyDeviceToHost));
  launch_column_dump(c->stream, c->nw, H, d0, diag, od, dv, cnt, ncap);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  unsigned long long m;
  HIPCHECK(c, hipMemcpy(&m, cnt, 8, hipMemcpyDeviceToHost));
  std::vector<D2> hd(m);
  std::vector<double> hv(m);
  HIPCHECK(c, hipMemcpy(hd.data(), od, m * sizeof(D2), hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(hv.data(), dv, m * 8, hipMemcpyDeviceToHost));
  for (void* p : {(void*)dd_, (void*)dh_, (void*)od, (void*)dv, (void*)cnt}) dfree(c, p);
  std::vector<size_t> order(m);
  for (size_t i = 0; i < m; ++i) order[i] = i;
  auto same = [&](const D2& a) { return a.lo == d0.lo && a.hi == d0.hi; };
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    bool da = same(hd[a]), db = same(hd[b]);
    if (da != db) return da;
    return det_less(hd[a], hd[b]);
  });
  *n = (int64_t)m;
  for (size_t i = 0; i < m && (int64_t)i < cap; ++i) {
    store_det(dets + i * c->nio, hd[order[i]], c->nio);
    vals[i] = hv[order[i]];
  }
  return MCDFCI_OK;
}

int mcdfci_submatrix(mcdfci_ctx* c, int k, const uint64_t* dets, double* Hout) {
  if (!c || !dets || !Hout || k < 1) return fail(c, MCDFCI_E_INPUT, "invalid argument");
  HIPCHECK(c, hipSetDevice(c->device));
  std::vector<D2> hd(k);
  int rc;
  for (int i = 0; i < k; ++i) {
    hd[i] = load_det(dets + (size_t)i * c->nio, c->nio);
    if ((rc = check_det(c, hd[i], "determinant"))) return rc;
  }
  D2* d = nullptr;
  double* h = nullptr;
  if ((rc = dalloc(c, &d, k))) return rc;
  if ((rc = dalloc(c, &h, (size_t)k * k))) return rc;
  HIPCHECK(c, hipMemcpy(d, hd.data(), sizeof(D2) * (size_t)k, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemsetAsync(h, 0, 8 * (size_t)k * k, c->stream));
  launch_submatrix(c->stream, c->nw, c->dh, nullptr, d, k, h, k);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(Hout, h, 8 * (size_t)k * k, hipMemcpyDeviceToHost));
  dfree(c, d);
  dfree(c, h);
  return MCDFCI_OK;
}

int mcdfci_eig_update(mcdfci_ctx* c, int k, const double* Hin, const double* x, const double* z, double* s,
                      double* dxo) {
  if (!c || !Hin || !x || !z || !s || !dxo || k < 1 || k > kMaxK) return fail(c, MCDFCI_E_INPUT, "invalid argument");
  HIPCHECK(c, hipSetDevice(c->device));
  double *h = nullptr, *dxx = nullptr, *dz = nullptr, *ddx = nullptr;
  int rc;
  if ((rc = dalloc(c, &h, (size_t)k * k))) return rc;
  if ((rc = dalloc(c, &dxx, k))) return rc;
  if ((rc = dalloc(c, &dz, k))) return rc;
  if ((rc = dalloc(c, &ddx, k))) return rc;
  // The update reads the upper triangle (j >= i); the kernel takes both.
  std::vector<double> hsym((size_t)k * k);
  for (int i = 0; i < k; ++i)
    for (int j = i; j < k; ++j) hsym[(size_t)i * k + j] = hsym[(size_t)j * k + i] = Hin[(size_t)i * k + j];
  HIPCHECK(c, hipMemcpy(h, hsym.data(), 8 * (size_t)k * k, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(dxx, x, 8 * (size_t)k, hipMemcpyHostToDevice));
  HIPCHECK(c, hipMemcpy(dz, z, 8 * (size_t)k, hipMemcpyHostToDevice));
  DevState st{};
  st.xx_hi = s[0];
  st.xx_lo = s[1];
  st.sc_hi = s[2];
  st.sc_lo = s[3];
  i128 xz = to_fix(s[4]) + to_fix(s[5]);
  st.xz_hi = (int64_t)(xz >> 64);
  st.xz_lo = (uint64_t)xz;
  HIPCHECK(c, hipMemcpy(c->st_scratch, &st, sizeof(st), hipMemcpyHostToDevice));
  launch_eig(c->stream, c->st_scratch, k, k, h, dxx, dz, ddx, nullptr, 0, nullptr, 1, c->xbuf);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(&st, c->st_scratch, sizeof(st), hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(dxo, ddx, 8 * (size_t)k, hipMemcpyDeviceToHost));
  for (void* p : {(void*)h, (void*)dxx, (void*)dz, (void*)ddx}) dfree(c, p);
  s[0] = st.xx_hi;
  s[1] = st.xx_lo;
  s[2] = st.sc_hi;
  s[3] = st.sc_lo;
  if (st.error == ERR_NUMERIC) return fail(c, MCDFCI_E_NUMERIC, "non-negative subspace eigenvalue");
  return MCDFCI_OK;
}

// mcdfci_eig_update without a device allocation: the inputs, the output dx
// and the solver state live in this device's `scratch`.  Every call on one
// GPU uses the same array (calls that overlap on a GPU race); each call
// writes all of what it reads before the kernels run.
int mcdfci_eig_update_noalloc(mcdfci_ctx* c, int k, const double* Hin, const double* x, const double* z,
                              double* s, double* dxo) {
  if (!c || !Hin || !x || !z || !s || !dxo) return fail(c, MCDFCI_E_INPUT, "null argument");
  if (k < 1 || k > kMaxK)
    return fail(c, MCDFCI_E_INPUT,
                "mcdfci_eig_update_noalloc: k must be in [1, 256] (its scratch holds a 256 x 256 subspace matrix)");
  if (!c->noalloc) return fail(c, MCDFCI_E_HIP, "no scratch address for this device");
  HIPCHECK(c, hipSetDevice(c->device));
  DevState* sd = (DevState*)c->noalloc;
  double* h = c->noalloc + kNoAllocState;
  double* xd = h + (size_t)k * k;
  double* zd = xd + k;
  double* dxd = zd + k;
  DevState st{};
  st.xx_hi = s[0];
  st.xx_lo = s[1];
  st.sc_hi = s[2];
  st.sc_lo = s[3];
  const i128 xz = to_fix(s[4]) + to_fix(s[5]);
  st.xz_hi = (int64_t)(xz >> 64);
  st.xz_lo = (uint64_t)xz;
  HIPCHECK(c, hipMemcpyAsync(sd, &st, sizeof(st), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(h, Hin, 8 * (size_t)k * k, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(xd, x, 8 * (size_t)k, hipMemcpyHostToDevice, c->stream));
  HIPCHECK(c, hipMemcpyAsync(zd, z, 8 * (size_t)k, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_mirror_upper, dim3((unsigned)std::min(64, (k * k + 255) / 256)), dim3(256), 0, c->stream, h,
                     k);
  launch_eig(c->stream, sd, k, k, h, xd, zd, dxd, nullptr, 0, nullptr, 1, (uint64_t*)(c->noalloc + kNoAllocXchg));
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipMemcpyAsync(&st, sd, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipMemcpyAsync(dxo, dxd, 8 * (size_t)k, hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  s[0] = st.xx_hi;
  s[1] = st.xx_lo;
  s[2] = st.sc_hi;
  s[3] = st.sc_lo;
  if (st.error == ERR_NUMERIC) return fail(c, MCDFCI_E_NUMERIC, "non-negative subspace eigenvalue");
  return MCDFCI_OK;
}

int mcdfci_table_dump(mcdfci_ctx* c, uint64_t* dets, double* x, double* z, int64_t cap, int64_t* n) {
  if (!c || !n) return fail(c, MCDFCI_E_INPUT, "null argument");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  DevState s;
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  long long ub = (long long)std::min<uint64_t>(c->nslots, (uint64_t)(s.z_size + s.x_size + 16 + c->k));
  uint64_t* od = nullptr;
  double* ox = nullptr;
  int64_t* oz = nullptr;
  unsigned long long* cnt = nullptr;
  int rc;
  if ((rc = dalloc(c, &od, 2 * ub))) return rc;
  if ((rc = dalloc(c, &ox, ub))) return rc;
  if ((rc = dalloc(c, &oz, 2 * ub))) return rc;
  if ((rc = dalloc(c, &cnt, 1))) return rc;
  HIPCHECK(c, hipMemsetAsync(cnt, 0, 8, c->stream));
  launch_table_compact(c->stream, c->nw, c->tab, c->nslots, od, ox, oz, cnt, ub);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  unsigned long long m;
  HIPCHECK(c, hipMemcpy(&m, cnt, 8, hipMemcpyDeviceToHost));
  if ((long long)m > ub) m = ub;
  std::vector<D2> hd(m);
  std::vector<double> hx(m);
  std::vector<int64_t> hz(2 * m);
  HIPCHECK(c, hipMemcpy(hd.data(), od, m * sizeof(D2), hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(hx.data(), ox, m * 8, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(hz.data(), oz, m * 16, hipMemcpyDeviceToHost));
  for (void* p : {(void*)od, (void*)ox, (void*)oz, (void*)cnt}) dfree(c, p);
  std::vector<size_t> order(m);
  for (size_t i = 0; i < m; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return det_less(hd[a], hd[b]); });
  *n = (int64_t)m;
  for (size_t i = 0; i < m && (int64_t)i < cap; ++i) {
    size_t j = order[i];
    if (dets) store_det(dets + i * c->nio, hd[j], c->nio);
    if (x) x[i] = hx[j];
    if (z) z[i] = fix_to_d(tz_join(hz[2 * j], hz[2 * j + 1]));
  }
  return MCDFCI_OK;
}

int mcdfci_table_digest(mcdfci_ctx* c, uint64_t* out3) {
  if (!c || !out3) return fail(c, MCDFCI_E_INPUT, "null argument");
  HIPCHECK(c, hipSetDevice(c->device));
  unsigned long long* acc = nullptr;
  int rc = dalloc(c, &acc, 3);
  if (rc) return rc;
  HIPCHECK(c, hipMemsetAsync(acc, 0, 24, c->stream));
  launch_table_digest(c->stream, c->nw, c->tab, c->nslots, acc);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  unsigned long long h[3];
  HIPCHECK(c, hipMemcpy(h, acc, 24, hipMemcpyDeviceToHost));
  dfree(c, acc);
  for (int i = 0; i < 3; ++i) out3[i] = h[i];
  return MCDFCI_OK;
}

int mcdfci_candidates(mcdfci_ctx* c, uint64_t* dets, double* x, double* z, int64_t cap, int64_t* n) {
  if (!c || !n) return fail(c, MCDFCI_E_INPUT, "null argument");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  DevState s;
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  std::vector<int> cnt(c->k);
  std::vector<CandEnt> ce((size_t)c->k * c->M);
  HIPCHECK(c, hipMemcpy(cnt.data(), c->cand_cnt, sizeof(int) * c->k, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(ce.data(), c->cand, sizeof(CandEnt) * ce.size(), hipMemcpyDeviceToHost));
  int64_t m = 0;
  for (int t = 0; t < s.k_prev; ++t)
    for (int i = 0; i < cnt[t]; ++i) {
      const CandEnt& e = ce[(size_t)t * c->M + i];
      // x, za, zb are the last three fields of either slot layout
      double xz[3];
      const size_t sb = slot_bytes(c->nw);
      HIPCHECK(c, hipMemcpy(xz, (const uint8_t*)c->tab + (size_t)e.slot * sb + sb - 24, 24, hipMemcpyDeviceToHost));
      int64_t za, zb;
      std::memcpy(&za, &xz[1], 8);
      std::memcpy(&zb, &xz[2], 8);
      if (m < cap) {
        if (dets) store_det(dets + m * c->nio, D2{e.det, e.det1}, c->nio);
        if (x) x[m] = xz[0];
        if (z) z[m] = fix_to_d(tz_join(za, zb));
      }
      ++m;
    }
  *n = m;
  return MCDFCI_OK;
}

}  // extern "C"

// FP64 Rayleigh quotient x.Hx / x.x of the table's x (fp64 integrals
// whatever the context's precision).
int mcdfci_energy_fp64(mcdfci_ctx* c, const mcdfci_ham* ham, double* energy) {
  if (!c || !energy) return fail(c, MCDFCI_E_INPUT, "null argument");
  if (!c->ready) return fail(c, MCDFCI_E_INPUT, "mcdfci_init has not been called");
  if (ham && (ham->norb != c->ham.norb || !ham->h || !ham->eri))
    return fail(c, MCDFCI_E_INPUT, "the integrals must have the context's NORB");
  if (ham && c->ham.uhf)
    return fail(c, MCDFCI_E_UNSUPPORTED, "caller integrals are spin-restricted; the context's are unrestricted");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  DevState s;
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  if (s.error) return check_state_error(c, s);
  long long ub = (long long)std::min<uint64_t>(c->nslots, (uint64_t)(s.x_size + c->k + 1024));
  D2* od = nullptr;
  double* ox = nullptr;
  unsigned long long* cnt = nullptr;
  int64_t* acc = nullptr;
  int rc;
  auto release = [&]() {
    for (void* p : {(void*)od, (void*)ox, (void*)cnt, (void*)acc}) dfree(c, p);
  };
  unsigned long long m = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if ((rc = dalloc(c, &od, ub)) || (rc = dalloc(c, &ox, ub)) || (rc = dalloc(c, &cnt, 1))) {
      release();
      return rc;
    }
    HIPCHECK(c, hipMemsetAsync(cnt, 0, 8, c->stream));
    launch_compact_x(c->stream, c->nw, c->tab, c->nslots, od, ox, cnt, ub);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&m, cnt, 8, hipMemcpyDeviceToHost));
    if ((long long)m <= ub) break;
    for (void* p : {(void*)od, (void*)ox, (void*)cnt}) dfree(c, p);
    od = nullptr, ox = nullptr, cnt = nullptr;
    ub = (long long)m;
  }
  if ((rc = dalloc(c, &acc, 4))) {
    release();
    return rc;
  }
  // Integrals: the context's, or the caller's (fp64 tables built here for
  // this call only).
  DevHam H = c->dh;
  std::vector<void*> tmp;
  auto release_tmp = [&]() {
    for (void* p : tmp) dfree(c, p);
    tmp.clear();
  };
  if (ham) {
    SpatialHam S;
    S.norb = c->ham.norb;
    S.core = ham->core;
    const size_t n2 = (size_t)S.norb * S.norb;
    S.h.assign(ham->h, ham->h + n2);
    S.eri.assign(ham->eri, ham->eri + n2 * n2);
    const Fp64Tables T = make_tables(S, c->opts.threshold);
    double *de = nullptr, *dh = nullptr, *dss = nullptr, *dso = nullptr, *dd1 = nullptr, *dd2 = nullptr;
    uint8_t* dsok = nullptr;
    if ((rc = upload(c, &de, S.eri)) || (rc = upload(c, &dh, S.h)) || (rc = upload(c, &dss, T.ss)) ||
        (rc = upload(c, &dso, T.so)) || (rc = upload(c, &dsok, T.sok)) || (rc = upload(c, &dd1, T.d1)) ||
        (rc = upload(c, &dd2, T.d2))) {
      for (void* p : {(void*)de, (void*)dh, (void*)dss, (void*)dso, (void*)dsok, (void*)dd1, (void*)dd2}) dfree(c, p);
      release();
      return rc;
    }
    tmp = {de, dh, dss, dso, dsok, dd1, dd2};
    H.core = S.core;
    H.eri = de;
    H.h1 = dh;
    H.s_same = dss;
    H.s_opp = dso;
    H.single_ok = dsok;
    H.d1 = dd1;
    H.d2 = dd2;
  }
  HIPCHECK(c, hipMemsetAsync(acc, 0, 32, c->stream));
  int64_t q[4];
  launch_sum_sq(c->stream, ox, (long long)m, 1.0, acc);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(q, acc, 16, hipMemcpyDeviceToHost));
  const double xx = fix_to_d(fix_join(q[0], q[1]));
  if (!(xx > 0.0)) {
    release();
    release_tmp();
    return fail(c, MCDFCI_E_NUMERIC, "the wavefunction is zero");
  }
  const double scale = 1.0 / std::sqrt(xx);
  HIPCHECK(c, hipMemsetAsync(acc, 0, 32, c->stream));
  launch_sum_sq(c->stream, ox, (long long)m, scale, acc);
  launch_rayleigh(c->stream, c->nw, H, c->tab, c->nslots - 1, od, ox, (long long)m, scale, acc + 2);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  HIPCHECK(c, hipMemcpy(q, acc, 32, hipMemcpyDeviceToHost));
  release();
  release_tmp();
  *energy = dd_div(fix_to_dd(fix_join(q[2], q[3])), fix_to_dd(fix_join(q[0], q[1]))).hi;
  return MCDFCI_OK;
}

// Gradient of the variational energy of the current wavefunction with
// respect to the unique integrals (see grad.hip).
int mcdfci_energy_gradient(mcdfci_ctx* c, double* grad_h, double* grad_eri, double* energy) {
  if (!c || !grad_h || !grad_eri) return fail(c, MCDFCI_E_INPUT, "null argument");
  if (!c->ready) return fail(c, MCDFCI_E_INPUT, "mcdfci_init has not been called");
  if (c->ham.uhf)
    return fail(c, MCDFCI_E_UNSUPPORTED, "the energy gradient is defined for spin-restricted integrals");
  HIPCHECK(c, hipSetDevice(c->device));
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  DevState s;
  HIPCHECK(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  if (s.error) return check_state_error(c, s);
  const int M = c->ham.norb;
  const size_t nh = (size_t)M * M, ne4 = nh * nh;
  long long ub = (long long)std::min<uint64_t>(c->nslots, (uint64_t)(s.x_size + c->k + 1024));
  D2* od = nullptr;
  double* ox = nullptr;
  unsigned long long* cnt = nullptr;
  int64_t *acc_h = nullptr, *acc_e = nullptr, *acc_q = nullptr;
  int rc;
  auto release = [&]() {
    for (void* p : {(void*)od, (void*)ox, (void*)cnt, (void*)acc_h, (void*)acc_e, (void*)acc_q}) dfree(c, p);
  };
  unsigned long long m = 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    if ((rc = dalloc(c, &od, ub)) || (rc = dalloc(c, &ox, ub)) || (rc = dalloc(c, &cnt, 1))) {
      release();
      return rc;
    }
    HIPCHECK(c, hipMemsetAsync(cnt, 0, 8, c->stream));
    launch_compact_x(c->stream, c->nw, c->tab, c->nslots, od, ox, cnt, ub);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemcpy(&m, cnt, 8, hipMemcpyDeviceToHost));
    if ((long long)m <= ub) break;
    for (void* p : {(void*)od, (void*)ox, (void*)cnt}) dfree(c, p);
    od = nullptr, ox = nullptr, cnt = nullptr;
    ub = (long long)m;
  }
  if ((rc = dalloc(c, &acc_h, 2 * nh)) || (rc = dalloc(c, &acc_e, 2 * ne4)) || (rc = dalloc(c, &acc_q, 2))) {
    release();
    return rc;
  }
  HIPCHECK(c, hipMemsetAsync(acc_h, 0, 16 * nh, c->stream));
  HIPCHECK(c, hipMemsetAsync(acc_e, 0, 16 * ne4, c->stream));
  HIPCHECK(c, hipMemsetAsync(acc_q, 0, 16, c->stream));
  int64_t q[2];
  // |x|^2 in fixed point, then the normalised sum c.c (exactly, for the
  // denominator of the Rayleigh quotient).
  launch_sum_sq(c->stream, ox, (long long)m, 1.0, acc_q);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));  // the copy below is on the null stream
  HIPCHECK(c, hipMemcpy(q, acc_q, 16, hipMemcpyDeviceToHost));
  const double xx = fix_to_d(fix_join(q[0], q[1]));
  if (!(xx > 0.0)) {
    release();
    return fail(c, MCDFCI_E_NUMERIC, "the wavefunction is zero");
  }
  const double scale = 1.0 / std::sqrt(xx);
  HIPCHECK(c, hipMemsetAsync(acc_q, 0, 16, c->stream));
  launch_sum_sq(c->stream, ox, (long long)m, scale, acc_q);
  launch_energy_grad(c->stream, c->nw, c->dh, c->tab, c->nslots - 1, od, ox, (long long)m, scale, acc_h, acc_e);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->stream));
  std::vector<int64_t> hh(2 * nh), he(2 * ne4);
  HIPCHECK(c, hipMemcpy(q, acc_q, 16, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(hh.data(), acc_h, 16 * nh, hipMemcpyDeviceToHost));
  HIPCHECK(c, hipMemcpy(he.data(), acc_e, 16 * ne4, hipMemcpyDeviceToHost));
  release();
  const dd den = fix_to_dd(fix_join(q[0], q[1]));
  double e = c->ham.core;
  for (size_t i = 0; i < nh; ++i) {
    grad_h[i] = dd_div(fix_to_dd(fix_join(hh[2 * i], hh[2 * i + 1])), den).hi;
    e += grad_h[i] * c->ham.h[i];
  }
  for (size_t i = 0; i < ne4; ++i) {
    grad_eri[i] = dd_div(fix_to_dd(fix_join(he[2 * i], he[2 * i + 1])), den).hi;
    e += grad_eri[i] * c->ham.eri[i];
  }
  if (energy) *energy = e;
  return MCDFCI_OK;
}
