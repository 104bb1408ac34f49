// Energy gradient with respect to the integrals (adjoint of the variational
// energy), computed from the wavefunction in the table.
//
// For the normalised coefficients c = x / |x| the variational energy is
//   E = sum_{D,D'} c_D c_D' H_{D'D} / sum_D c_D^2,
// and H_{D'D} is linear in the integrals, so at fixed c
//   dE/dtheta = sum_{D,D'} c_D c_D' dH_{D'D}/dtheta / sum_D c_D^2
// for every unique FCIDUMP parameter theta (h_pq with p >= q; (pq|rs) in its
// canonical 8-fold position).  This equals contracting the spin-summed
// 1-RDM and 2-RDM with the integral derivatives (Hellmann-Feynman at an
// eigenvector).  Every ordered pair (D, D') with D' a single or double
// excitation of D (or D itself) contributes the Slater-Condon terms of
// H_{D'D} (src/hamiltonian.h:257-410), with no threshold filter: a parameter
// that is zero now still has a derivative.  The sums are exact fixed-point
// integers (2^-64 units), so the result does not depend on the order in which
// the pairs are visited.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "column.h"
#include "device.h"
#include "grad.h"

namespace mc {

// Canonical positions of the unique parameters.
__device__ __forceinline__ int h_index(int M, int p, int q) { return p >= q ? p * M + q : q * M + p; }

__device__ __forceinline__ size_t eri_index(int M, int p, int q, int r, int s) {
  int a = p >= q ? p : q, b = p >= q ? q : p;
  int c = r >= s ? r : s, d = r >= s ? s : r;
  const int P1 = a * (a + 1) / 2 + b, P2 = c * (c + 1) / 2 + d;
  if (P1 < P2) {
    int t = a;
    a = c;
    c = t;
    t = b;
    b = d;
    d = t;
  }
  return (((size_t)a * M + b) * M + c) * M + d;
}

__device__ __forceinline__ void acc_add(int64_t* acc, size_t idx, int64_t a, int64_t b, bool neg) {
  if (neg) {
    a = -a;
    b = -b;
  }
  if (a) __hip_atomic_fetch_add(&acc[2 * idx], a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (b) __hip_atomic_fetch_add(&acc[2 * idx + 1], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Determinants with x != 0 (any order: the sums below are exact).
template <int W>
__global__ void k_compact_x(const SlotW<W>* tab, uint64_t nslots, D2* odet, double* ox, unsigned long long* count,
                            long long cap) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += stride) {
    if (i == 0xffffffffULL) continue;
    const SlotW<W> s = tab[i];
    D2 d;
    if constexpr (W == 1) {
      if (s.key == 0) continue;
      d = D2{s.key & kKeyMask, 0};
    } else {
      if (s.k1 == 0) continue;
      d = D2{s.k0, s.k1 & kHiMask};
    }
    if (!(__builtin_fabs(s.x) != 0.0)) continue;
    const unsigned long long pos = atomicAdd(count, 1ULL);
    if ((long long)pos < cap) {
      odet[pos] = d;
      ox[pos] = s.x;
    }
  }
}

// sum x^2 in fixed point.
__global__ void k_sum_sq(const double* x, long long n, double scale, int64_t* acc) {
  i128 v = 0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double c = x[i] * scale;
    v += to_fix(c * c);
  }
  int64_t a, b;
  fix_split(v, a, b);
  // Partial sums of a and b stay far inside int64 (|c| <= 1 after scaling).
  for (int off = 32; off >= 1; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  if ((threadIdx.x & 63) == 0) acc_add(acc, 0, a, b, false);
}

// One workgroup per (determinant, column chunk); the table supplies x of
// the connected determinant.
template <int W>
__global__ __launch_bounds__(kWG) void k_energy_grad(DevHam H, const SlotW<W>* tab, uint64_t mask, const D2* dets,
                                                     const double* xs, long long n, double scale, int64_t* acc_h,
                                                     int64_t* acc_e) {
  __shared__ ColCtxW<W> C;
  const int M = H.M;
  const int chunk = blockIdx.x;
  for (long long di = blockIdx.y; di < n; di += gridDim.y) {
    setup_ctx<W>(C, H, dets[di]);
    const D2 D = C.det;
    const double cD = xs[di] * scale;
    if (chunk == 0) {
      // Diagonal: h_II for occupied I, (II|JJ) - delta_spin (IJ|JI) for I < J.
      int64_t wa, wb;
      fix_split(to_fix(cD * cD), wa, wb);
      const int ne = C.ne;
      const int npair = ne * (ne - 1) / 2;
      for (int t = threadIdx.x; t < ne + npair; t += blockDim.x) {
        if (t < ne) {
          const int p = C.occ[t] >> 1;
          acc_add(acc_h, h_index(M, p, p), wa, wb, false);
        } else {
          int f = t - ne, x = 0;
          while (f >= ne - 1 - x) {
            f -= ne - 1 - x;
            ++x;
          }
          const int y = x + 1 + f;
          const int I = C.occ[x], J = C.occ[y];
          const int p = I >> 1, q = J >> 1;
          acc_add(acc_e, eri_index(M, p, p, q, q), wa, wb, false);
          if ((I & 1) == (J & 1)) acc_add(acc_e, eri_index(M, p, q, q, p), wa, wb, true);
        }
      }
    }
    for (int r = 0; r < kEPT; ++r) {
      const int e = chunk * kChunk + r * kWG + threadIdx.x;
      if (e == 0 || e >= H.n_enum) continue;
      // Decode as in the column generator, without the threshold filter.
      int kind;  // 0 single, 1 same-spin double, 2 alpha-beta double
      int p, q = 0, rr, s = 0, spin = 0;
      int i, j = 0, a, b = 0;
      if (e < H.off_daa) {
        int f = e - 1, o, u;
        if (f < H.n_sa) {
          spin = 0;
          o = f / H.nva;
          u = f - o * H.nva;
        } else {
          spin = 1;
          f -= H.n_sa;
          o = f / H.nvb;
          u = f - o * H.nvb;
        }
        p = spin ? C.ob[o] : C.oa[o];
        rr = spin ? C.vb[u] : C.va[u];
        kind = 0;
        i = 2 * p + spin;
        a = 2 * rr + spin;
      } else if (e < H.off_dab) {
        int f, cv;
        if (e < H.off_dbb) {
          spin = 0;
          f = e - H.off_daa;
          cv = H.cpa_v;
        } else {
          spin = 1;
          f = e - H.off_dbb;
          cv = H.cpb_v;
        }
        const int op = f / cv, vp = f - op * cv;
        const uint16_t po = spin ? C.pob[op] : C.poa[op];
        const uint16_t pv = spin ? C.pvb[vp] : C.pva[vp];
        const uint8_t* ol = spin ? C.ob : C.oa;
        const uint8_t* vl = spin ? C.vb : C.va;
        p = ol[po & 0xff];
        q = ol[po >> 8];
        rr = vl[pv & 0xff];
        s = vl[pv >> 8];
        kind = 1;
        i = 2 * p + spin;
        j = 2 * q + spin;
        a = 2 * rr + spin;
        b = 2 * s + spin;
      } else {
        const int f = e - H.off_dab;
        const int vbi = f % H.nvb;
        int t = f / H.nvb;
        const int vai = t % H.nva;
        t = t / H.nva;
        const int obi = t % H.nb;
        const int oai = t / H.nb;
        p = C.oa[oai];
        q = C.ob[obi];
        rr = C.va[vai];
        s = C.vb[vbi];
        kind = 2;
        if (p <= q) {
          i = 2 * p;
          j = 2 * q + 1;
          a = 2 * rr;
          b = 2 * s + 1;
        } else {
          i = 2 * q + 1;
          j = 2 * p;
          a = 2 * s + 1;
          b = 2 * rr;
        }
      }
      D2 d1 = D;
      dflip<W>(d1, i);
      int sg = dparity<W>(d1, i, a);
      dflip<W>(d1, a);
      if (kind != 0) {
        dflip<W>(d1, j);
        sg ^= dparity<W>(d1, j, b);
        dflip<W>(d1, b);
      }
      const long long si = find_slot<W>(tab, mask, d1);
      if (si < 0) continue;
      const double xo = tab[si].x;
      if (!(__builtin_fabs(xo) != 0.0)) continue;
      int64_t wa, wb;
      fix_split(to_fix(cD * (xo * scale)), wa, wb);
      const bool neg = sg != 0;
      if (kind == 0) {
        // h_pr + sum_{K in occ(D)} [(pr|KK) - delta_spin (pK|Kr)]
        acc_add(acc_h, h_index(M, p, rr), wa, wb, neg);
        for (int x = 0; x < C.ne; ++x) {
          const int K = C.occ[x], k = K >> 1;
          acc_add(acc_e, eri_index(M, p, rr, k, k), wa, wb, neg);
          if ((K & 1) == spin) acc_add(acc_e, eri_index(M, p, k, k, rr), wa, wb, !neg);
        }
      } else if (kind == 1) {
        // (pr|qs) - (ps|qr)
        acc_add(acc_e, eri_index(M, p, rr, q, s), wa, wb, neg);
        acc_add(acc_e, eri_index(M, p, s, q, rr), wa, wb, !neg);
      } else {
        acc_add(acc_e, eri_index(M, p, rr, q, s), wa, wb, neg);
      }
    }
    __syncthreads();  // C is rebuilt for the next determinant
  }
}

void launch_compact_x(hipStream_t s, int nw, const void* tab, uint64_t nslots, D2* odet, double* ox,
                      unsigned long long* count, long long cap) {
  const uint64_t blocks = std::min<uint64_t>((nslots + 255) / 256, kSlotBlocks);
  if (nw == 1)
    hipLaunchKernelGGL(k_compact_x<1>, dim3((unsigned)blocks), dim3(256), 0, s, (const SlotW<1>*)tab, nslots, odet,
                       ox, count, cap);
  else
    hipLaunchKernelGGL(k_compact_x<2>, dim3((unsigned)blocks), dim3(256), 0, s, (const SlotW<2>*)tab, nslots, odet,
                       ox, count, cap);
}

void launch_sum_sq(hipStream_t s, const double* x, long long n, double scale, int64_t* acc) {
  hipLaunchKernelGGL(k_sum_sq, dim3(256), dim3(256), 0, s, x, n, scale, acc);
}

void launch_energy_grad(hipStream_t s, int nw, const DevHam& H, const void* tab, uint64_t mask, const D2* dets,
                        const double* xs, long long n, double scale, int64_t* acc_h, int64_t* acc_e) {
  if (n <= 0) return;
  const int nchunks = (H.n_enum + kChunk - 1) / kChunk;
  const long long gy = n < 65535 ? n : 65535;
  if (nw == 1)
    hipLaunchKernelGGL(k_energy_grad<1>, dim3(nchunks, (unsigned)gy), dim3(kWG), 0, s, H, (const SlotW<1>*)tab,
                       mask, dets, xs, n, scale, acc_h, acc_e);
  else
    hipLaunchKernelGGL(k_energy_grad<2>, dim3(nchunks, (unsigned)gy), dim3(kWG), 0, s, H, (const SlotW<2>*)tab,
                       mask, dets, xs, n, scale, acc_h, acc_e);
}

}  // namespace mc
