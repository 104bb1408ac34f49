// Test-only shim around the reference's own headers (compiled from the
// sources where they lie, never copied).  It exposes, through a C ABI that
// the pytest harness loads with ctypes, the parts of the reference that
// build without external libraries:
//   * Fcidump parsing                       (src/input.h:57-338)
//   * HamiltonianMolecule<N> tables, get_column, get_diagonal,
//     get_off_diagonal, get_hartree_fock    (src/hamiltonian.h:56-493)
//   * CDFCI<N>::get_hamiltonian_submatrix   (src/solver.h:637-670)
//   * CoordinatePickGcdGrad<N>              (src/solver.h:106-225)
//   * WaveFunctionCuckoo<N>::update_x / update_z_and_get_sub / fetch_val
//                                           (src/wavefunction.h:398-668)
//   * CoordinateUpdateLS<N>                 (src/solver.h:254-314)
// CoordinateUpdateEig (src/solver.h:319-610) needs LAPACK/BLAS, which this
// image does not provide, so it is never instantiated here.
//
// This file is oracle infrastructure: it is only loaded by tests/.

#include "solver.h"

#include <cstdint>
#include <cstring>
#include <string>

namespace {

std::string g_err;

struct HamBase {
  int nw = 1;
  int norb = 0;  // spin-orbitals
  int nelec = 0;
  int ms2 = 0;
  virtual ~HamBase() {}
};

template <int N>
struct HamN : HamBase {
  Fcidump fci;
  HamiltonianMolecule<N> h;
  HamN(const std::string& path, double thr) : fci(path), h(fci, thr) {
    nw = N;
    norb = fci.norb;
    nelec = fci.nelec;
    ms2 = fci.ms2;
  }
};

template <int N>
Determinant<N> to_det(const uint64_t* w) {
  Determinant<N> d;
  for (int i = 0; i < N; ++i) d.repr[i] = (size_t)w[i];
  return d;
}

template <int N>
void from_det(const Determinant<N>& d, uint64_t* w) {
  for (int i = 0; i < N; ++i) w[i] = (uint64_t)d.repr[i];
}

struct WfBase {
  int nw = 1;
  virtual ~WfBase() {}
};

template <int N>
struct WfN : WfBase {
  WaveFunctionCuckoo<N> wf;
  WfN(long cap, int mm) : wf((size_t)cap) {
    nw = N;
    wf.MM = mm;
  }
};

}  // namespace

extern "C" const char* ref_last_error() { return g_err.c_str(); }

extern "C" void* ref_ham_create(const char* path, double threshold) {
  try {
    // Word count rule of run_wrapper.h:181.
    Fcidump probe(path);
    int n = 1 + (probe.norb - 1) / 64;
    Determinant<1>::constuct_masks();
    Determinant<2>::constuct_masks();
    if (n == 1) return new HamN<1>(path, threshold);
    if (n == 2) return new HamN<2>(path, threshold);
    g_err = "only 1 or 2 words per determinant are exposed by the shim";
    return nullptr;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

extern "C" void ref_ham_destroy(void* p) { delete (HamBase*)p; }

extern "C" void ref_ham_info(void* p, int* out) {
  HamBase* h = (HamBase*)p;
  out[0] = h->nw;
  out[1] = h->norb;
  out[2] = h->nelec;
  out[3] = h->ms2;
}

extern "C" void ref_hf(void* p, uint64_t* det) {
  HamBase* b = (HamBase*)p;
  if (b->nw == 1) from_det<1>(((HamN<1>*)b)->h.get_hartree_fock(), det);
  else from_det<2>(((HamN<2>*)b)->h.get_hartree_fock(), det);
}

template <int N>
static double diag_impl(HamN<N>* h, const uint64_t* det) {
  auto d = to_det<N>(det);
  DeterminantDecoded<N> dd(d);
  return h->h.get_diagonal(dd);
}

extern "C" double ref_diagonal(void* p, const uint64_t* det) {
  HamBase* b = (HamBase*)p;
  return b->nw == 1 ? diag_impl<1>((HamN<1>*)b, det) : diag_impl<2>((HamN<2>*)b, det);
}

template <int N>
static double offdiag_impl(HamN<N>* h, const uint64_t* d1, const uint64_t* d2) {
  auto a = to_det<N>(d1);
  auto b = to_det<N>(d2);
  DeterminantDecoded<N> da(a), db(b);
  return h->h.get_off_diagonal(da, db);
}

extern "C" double ref_off_diagonal(void* p, const uint64_t* d1, const uint64_t* d2) {
  HamBase* b = (HamBase*)p;
  return b->nw == 1 ? offdiag_impl<1>((HamN<1>*)b, d1, d2)
                    : offdiag_impl<2>((HamN<2>*)b, d1, d2);
}

template <int N>
static long column_impl(HamN<N>* h, const uint64_t* det, uint64_t* dets, double* vals, long cap) {
  auto d = to_det<N>(det);
  DeterminantDecoded<N> dd(d);
  auto col = h->h.get_column(dd);
  long n = (long)col.size();
  for (long i = 0; i < n && i < cap; ++i) {
    from_det<N>(col[i].first, dets + i * N);
    vals[i] = col[i].second;
  }
  return n;
}

extern "C" long ref_column(void* p, const uint64_t* det, uint64_t* dets, double* vals, long cap) {
  HamBase* b = (HamBase*)p;
  return b->nw == 1 ? column_impl<1>((HamN<1>*)b, det, dets, vals, cap)
                    : column_impl<2>((HamN<2>*)b, det, dets, vals, cap);
}

template <int N>
static void submatrix_impl(HamN<N>* h, int k, const uint64_t* dets, double* H) {
  CDFCI<N> solver;
  WaveFunctionVector<N> picked;
  for (int i = 0; i < k; ++i) picked.data.push_back({to_det<N>(dets + i * N), {0.0, 0.0}});
  solver.get_hamiltonian_submatrix(h->h, picked, H);
}

extern "C" void ref_submatrix(void* p, int k, const uint64_t* dets, double* H) {
  HamBase* b = (HamBase*)p;
  if (b->nw == 1) submatrix_impl<1>((HamN<1>*)b, k, dets, H);
  else submatrix_impl<2>((HamN<2>*)b, k, dets, H);
}

template <int N>
static void pick_impl(long n, const uint64_t* dets, const double* x, const double* z, double xx,
                      double scale, int k, uint64_t* out_dets, double* out_x, double* out_z) {
  WaveFunctionVector<N> list;
  for (long i = 0; i < n; ++i) list.data.push_back({to_det<N>(dets + i * N), {x[i], z[i]}});
  list.set_xx(xx);
  list.set_scale(scale);
  CoordinatePickGcdGrad<N> pick;
  auto res = pick(list, k);
  for (int i = 0; i < k; ++i) {
    from_det<N>(res.data[i].first, out_dets + i * N);
    out_x[i] = res.data[i].second[0];
    out_z[i] = res.data[i].second[1];
  }
}

extern "C" void ref_pick(int nw, long n, const uint64_t* dets, const double* x, const double* z, double xx,
              double scale, int k, uint64_t* out_dets, double* out_x, double* out_z) {
  if (nw == 1) pick_impl<1>(n, dets, x, z, xx, scale, k, out_dets, out_x, out_z);
  else pick_impl<2>(n, dets, x, z, xx, scale, k, out_dets, out_x, out_z);
}

extern "C" void* ref_wf_create(int nw, long capacity, int mm) {
  if (nw == 1) return new WfN<1>(capacity, mm);
  return new WfN<2>(capacity, mm);
}

extern "C" void ref_wf_destroy(void* p) { delete (WfBase*)p; }

template <int N>
static void wf_scalars_impl(WfN<N>* w, int set, double* s) {
  if (set) {
    w->wf.set_xx(s[0]);
    w->wf.set_scale(s[1]);
    w->wf.set_xz(s[2]);
  } else {
    s[0] = (double)w->wf.xx();
    s[1] = (double)w->wf.scale();
    s[2] = (double)w->wf.xz();
    s[3] = (double)w->wf.x_size();
    s[4] = (double)w->wf.size();
    s[5] = (double)w->wf.updates_size();
    s[6] = w->wf.get_variational_energy();
  }
}

// set != 0: s = {xx, scale, xz}; set == 0: s[0..6] = {xx, scale, xz,
// x_size, z_size, updates, energy}.
extern "C" void ref_wf_scalars(void* p, int set, double* s) {
  WfBase* b = (WfBase*)p;
  if (b->nw == 1) wf_scalars_impl<1>((WfN<1>*)b, set, s);
  else wf_scalars_impl<2>((WfN<2>*)b, set, s);
}

template <int N>
static void wf_update_x_impl(WfN<N>* w, int k, const uint64_t* dets, const double* dx) {
  WaveFunctionVector<N> picked;
  for (int i = 0; i < k; ++i) picked.data.push_back({to_det<N>(dets + i * N), {0.0, 0.0}});
  std::vector<double> d(dx, dx + k);
  w->wf.update_x(picked, d);
}

extern "C" void ref_wf_update_x(void* p, int k, const uint64_t* dets, const double* dx) {
  WfBase* b = (WfBase*)p;
  if (b->nw == 1) wf_update_x_impl<1>((WfN<1>*)b, k, dets, dx);
  else wf_update_x_impl<2>((WfN<2>*)b, k, dets, dx);
}

// Runs update_z_and_get_sub for k columns given as CSR (col_off[k+1]).
// Writes the candidate list; returns its length (or -1 on overflow error).
template <int N>
static long wf_update_z_impl(WfN<N>* w, int k, const long* col_off, const uint64_t* col_dets,
                             const double* col_vals, const uint64_t* picked_dets, const double* dx,
                             double z_thr, uint64_t* cand_dets, double* cand_x, double* cand_z,
                             long cap) {
  std::vector<std::vector<std::pair<Determinant<N>, double>>> cols(k);
  for (int t = 0; t < k; ++t)
    for (long e = col_off[t]; e < col_off[t + 1]; ++e)
      cols[t].push_back({to_det<N>(col_dets + e * N), col_vals[e]});
  WaveFunctionVector<N> picked, sub;
  for (int i = 0; i < k; ++i) picked.data.push_back({to_det<N>(picked_dets + i * N), {0.0, 0.0}});
  std::vector<double> d(dx, dx + k);
  try {
    w->wf.update_z_and_get_sub(cols.data(), picked, d, sub, z_thr);
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  long n = (long)sub.data.size();
  for (long i = 0; i < n && i < cap; ++i) {
    from_det<N>(sub.data[i].first, cand_dets + i * N);
    cand_x[i] = sub.data[i].second[0];
    cand_z[i] = sub.data[i].second[1];
  }
  return n;
}

extern "C" long ref_wf_update_z(void* p, int k, const long* col_off, const uint64_t* col_dets,
                     const double* col_vals, const uint64_t* picked_dets, const double* dx,
                     double z_thr, uint64_t* cand_dets, double* cand_x, double* cand_z, long cap) {
  WfBase* b = (WfBase*)p;
  if (b->nw == 1)
    return wf_update_z_impl<1>((WfN<1>*)b, k, col_off, col_dets, col_vals, picked_dets, dx, z_thr,
                               cand_dets, cand_x, cand_z, cap);
  return wf_update_z_impl<2>((WfN<2>*)b, k, col_off, col_dets, col_vals, picked_dets, dx, z_thr,
                             cand_dets, cand_x, cand_z, cap);
}

extern "C" int ref_wf_fetch(void* p, const uint64_t* det, double* x, double* z) {
  WfBase* b = (WfBase*)p;
  if (b->nw == 1) {
    auto d = to_det<1>(det);
    ((WfN<1>*)b)->wf.fetch_val(d, 0.0, *x, *z);
  } else {
    auto d = to_det<2>(det);
    ((WfN<2>*)b)->wf.fetch_val(d, 0.0, *x, *z);
  }
  return 0;
}

// CoordinateUpdateLS::operator() on one picked coordinate (x, z) with
// diagonal h and |x|^2 = xx held by the wavefunction.
extern "C" double ref_ls_update(double h, double x, double z, double xx) {
  WfN<1> w(64, 1);
  w.wf.set_xx(xx);
  WaveFunctionVector<1> sub;
  sub.data.push_back({Determinant<1>(), {x, z}});
  double H[1] = {h};
  CoordinateUpdateLS<1> ls;
  return ls(H, w.wf, sub)[0];
}
