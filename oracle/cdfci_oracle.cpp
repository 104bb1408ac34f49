// CPU oracle for the CDFCI hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: it
// is loaded by tests/ (ctypes), by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, and nothing in the product links it.
//
// It restates the reference's CPU algorithm:
//   * FCIDUMP dialect and RHF spin expansion      src/input.h:93-330
//   * excitation tables, get_diagonal, singles,
//     doubles, get_column, get_off_diagonal, HF   src/hamiltonian.h:98-492
//   * determinant parity                          src/determinant.h:102-111,148-152
//   * gcd_grad pick                               src/solver.h:196-224
//   * eig update                                  src/solver.h:333-609 (oracle_eig.h)
//   * update_x / update_z_and_get_sub             src/wavefunction.h:425-605
//   * the iteration loop and report               src/solver.h:672-893
// with the reference's thread races (k > 1) fixed to one canonical order:
//   - z is an exact fixed-point accumulator (2^-64 units), so the result of
//     the scatter-add does not depend on the order of the adds;
//   - an element whose dz fails the compression gate updates a determinant
//     only if that determinant was in the table before this iteration;
//   - each column's top-M list ranks by (|x*xss+z| desc, determinant desc)
//     using the values after all columns' adds and before the z reinsert;
//   - new_z, xz and the per-column dot products are exact fixed-point sums;
//   - xx, scale and xz are double-double / fixed point instead of binary128.
//
// Determinants are the reference's interleaved bit strings (bit 2p = alpha
// of spatial orbital p, bit 2p+1 = beta), one or two 64-bit words.

#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "oracle_eig.h"
#include "oracle_ls.h"
#include "oracle_num.h"

namespace oc {

struct Det {
  uint64_t w[2] = {0, 0};
  bool operator==(const Det& o) const { return w[0] == o.w[0] && w[1] == o.w[1]; }
  bool operator!=(const Det& o) const { return !(*this == o); }
  bool occ(int s) const { return (w[s >> 6] >> (s & 63)) & 1; }
  void set(int s) { w[s >> 6] |= (uint64_t)1 << (s & 63); }
  void clr(int s) { w[s >> 6] &= ~((uint64_t)1 << (s & 63)); }
};

// Descending determinant order used for ties: compare from word 0, unsigned
// (src/wavefunction.h:30-36 compares from word 0 by the sign of the
// difference; the two agree whenever the words differ by less than 2^63).
static inline bool det_greater(const Det& a, const Det& b) {
  if (a.w[0] != b.w[0]) return a.w[0] > b.w[0];
  return a.w[1] > b.w[1];
}

static inline uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}
static inline uint64_t det_hash(const Det& d) { return mix64(d.w[0] ^ mix64(d.w[1] + 0x9e3779b97f4a7c15ULL)); }

// Parity of occupied spin-orbitals in [min(i,a), max(i,a)) (determinant.h:102-111).
static inline int parity(const Det& d, int i, int a) {
  int lo = std::min(i, a), hi = std::max(i, a);
  int c = 0;
  for (int w = 0; w < 2; ++w) {
    int b0 = w * 64, b1 = b0 + 64;
    int l = std::max(lo, b0), h = std::min(hi, b1);
    if (l >= h) continue;
    uint64_t m = (h - b0 == 64 ? ~0ULL : ((1ULL << (h - b0)) - 1)) & ~((1ULL << (l - b0)) - 1);
    c += __builtin_popcountll(d.w[w] & m);
  }
  return c & 1;
}

static std::vector<int> occupied(const Det& d, int nbits) {
  std::vector<int> o;
  for (int s = 0; s < nbits; ++s)
    if (d.occ(s)) o.push_back(s);
  return o;
}

// ---------------------------------------------------------------- FCIDUMP
struct Fcidump {
  int norb = 0;  // spin-orbitals
  int nelec = 0, ms2 = 0;
  bool uhf = false;
  double core = 0.0;
  std::vector<double> h1;                     // norb x norb
  std::unordered_map<uint64_t, double> g;     // <ij|ab>, i <= j keys, first insert wins
  std::vector<std::pair<int, double>> orb_e;  // first insert wins per orbital
  bool has_orb_e = false;

  static uint64_t key(int i, int j, int a, int b) {
    return ((uint64_t)i << 48) | ((uint64_t)j << 32) | ((uint64_t)a << 16) | (uint64_t)b;
  }
  void ins(int i, int j, int a, int b, double v) {
    if (i <= j) g.emplace(key(i, j, a, b), v);
    if (i >= j) g.emplace(key(j, i, b, a), v);
  }
  double get2(int i, int j, int a, int b) const {
    auto it = g.find(key(i, j, a, b));
    return it == g.end() ? 0.0 : it->second;
  }
  void add_orb_e(int i, double v) {
    for (auto& p : orb_e)
      if (p.first == i) return;
    orb_e.push_back({i, v});
    has_orb_e = true;
  }

  static int read_int(const std::string& header, const std::string& name) {
    std::regex r(name + R"([ ]*=[ ]*(\d+))");
    std::smatch m;
    if (!std::regex_search(header, m, r)) throw std::invalid_argument(name + " is not found.");
    return std::stoi(m[1]);
  }

  void load(const std::string& path) {
    std::ifstream f(path);
    if (f.fail()) throw std::invalid_argument("Failed to open the FCIDUMP file " + path);
    std::string header, line;
    while (std::getline(f, line) && line.find("&END") == std::string::npos &&
           line.find("/") == std::string::npos) {
      header += " ";
      header += line;
    }
    header += " &END";
    if (f.eof()) throw std::invalid_argument("FCIDUMP has the wrong header.");
    norb = read_int(header, "NORB");
    nelec = read_int(header, "NELEC");
    ms2 = read_int(header, "MS2");
    {
      std::regex r(R"(UHF[ ]*=[ .]*(FALSE|TRUE))");
      std::smatch m;
      uhf = std::regex_search(header, m, r) && m[1] == "TRUE";
    }
    if (!uhf) norb *= 2;
    h1.assign((size_t)norb * norb, 0.0);
    double v;
    int i, a, j, b;
    while (f >> v >> i >> a >> j >> b) {
      if (i && a && j && b) {
        if (uhf) {
          ins(i - 1, j - 1, a - 1, b - 1, v);
          ins(i - 1, b - 1, a - 1, j - 1, v);
          ins(a - 1, j - 1, i - 1, b - 1, v);
          ins(a - 1, b - 1, i - 1, j - 1, v);
        } else {
          rhf4(i, j, a, b, v);
          rhf4(i, b, a, j, v);
          rhf4(a, j, i, b, v);
          rhf4(a, b, i, j, v);
        }
      } else if (a) {
        if (uhf) {
          h1[(i - 1) * norb + (a - 1)] = v;
          h1[(a - 1) * norb + (i - 1)] = v;
        } else {
          int ia = 2 * i - 2, ib = 2 * i - 1, aa = 2 * a - 2, ab = 2 * a - 1;
          h1[ia * norb + aa] = v;
          h1[ib * norb + ab] = v;
          h1[aa * norb + ia] = v;
          h1[ab * norb + ib] = v;
        }
      } else if (i) {
        if (uhf) {
          add_orb_e(i - 1, v);
        } else {
          add_orb_e(2 * i - 2, v);
          add_orb_e(2 * i - 1, v);
        }
      } else {
        core = v;
      }
    }
  }
  void rhf4(int i, int j, int a, int b, double v) {
    auto al = [](int x) { return 2 * x - 2; };
    auto be = [](int x) { return 2 * x - 1; };
    ins(al(i), al(j), al(a), al(b), v);
    ins(be(i), be(j), be(a), be(b), v);
    ins(al(i), be(j), al(a), be(b), v);
    ins(be(i), al(j), be(a), al(b), v);
  }
};

// ---------------------------------------------------------------- tables
struct Ham {
  int ns = 0;  // spin-orbitals
  int nelec = 0, ms2 = 0, nw = 1;
  double core = 0.0;
  // Mixed precision (the device's precision = mixed): column elements other
  // than the diagonal are evaluated in fp32 from fp32 roundings of the fp64
  // integrals, in the device's operation order; the diagonal, the
  // off-diagonal elements of the subspace matrix and everything downstream
  // stay fp64.
  bool mixed = false;
  struct D2 {
    int a, b;
    double v;
    float v32;         // fp32 value: (float)<ij|ab> - (float)<ij|ba> or (float)<ij|ab>
    bool ok64, ok32;   // |v| > threshold, |v32| > threshold
  };
  std::vector<std::vector<D2>> dbl;  // ns*ns, i < j
  struct S1 {
    int a;
    double h;
    std::vector<double> two;  // <ik||ak> for all k
  };
  std::vector<std::vector<S1>> sgl;
  std::vector<double> d1, d2;
  std::vector<std::pair<int, double>> orb_e;
  bool has_orb_e = false;

  void build(const Fcidump& f, double thr) {
    ns = f.norb;
    nelec = f.nelec;
    ms2 = f.ms2;
    core = f.core;
    nw = 1 + (ns - 1) / 64;
    orb_e = f.orb_e;
    has_orb_e = f.has_orb_e;
    // Doubles (hamiltonian.h:115-173).
    dbl.assign((size_t)ns * ns, {});
    for (auto& kv : f.g) {
      int i = (int)(kv.first >> 48) & 0xffff, j = (int)(kv.first >> 32) & 0xffff;
      int a = (int)(kv.first >> 16) & 0xffff, b = (int)kv.first & 0xffff;
      double v = kv.second;
      if (i == j) continue;
      // An entry is kept when its fp64 value or its fp32 value passes the
      // strict threshold; each precision's column uses its own flag.
      auto keep = [&](int ea, int eb, double w, float w32) {
        D2 e{ea, eb, w, w32, std::fabs(w) > thr, std::fabs((double)w32) > thr};
        if (e.ok64 || e.ok32) dbl[i * ns + j].push_back(e);
      };
      if ((i % 2) ^ (j % 2)) {
        keep(a, b, v, (float)v);
      } else if (a < b) {
        const double u = f.get2(i, j, b, a);
        keep(a, b, v - u, (float)v - (float)u);
      } else if (b < a && f.g.find(Fcidump::key(i, j, b, a)) == f.g.end()) {
        keep(b, a, -v, -(float)v);
      }
    }
    for (auto& l : dbl) {
      std::stable_sort(l.begin(), l.end(), [](const D2& x, const D2& y) {
        if (std::fabs(x.v) != std::fabs(y.v)) return std::fabs(x.v) > std::fabs(y.v);
        if (x.a != y.a) return x.a < y.a;
        return x.b < y.b;
      });
    }
    // Singles (hamiltonian.h:176-219).
    sgl.assign(ns, {});
    for (int i = 0; i < ns; ++i) {
      for (int a = 0; a < ns; ++a) {
        S1 e;
        e.a = a;
        e.h = f.h1[i * ns + a];
        e.two.resize(ns);
        double mx = std::fabs(e.h);
        for (int k = 0; k < ns; ++k) {
          double v;
          if (i <= k) v = f.get2(i, k, a, k) - f.get2(i, k, k, a);
          else v = f.get2(k, i, k, a) - f.get2(k, i, a, k);
          e.two[k] = v;
          mx = std::max(mx, std::fabs(v));
        }
        if (mx > thr) sgl[i].push_back(e);
      }
      std::stable_sort(sgl[i].begin(), sgl[i].end(),
                       [](const S1& x, const S1& y) { return std::fabs(x.h) > std::fabs(y.h); });
    }
    // Diagonal (hamiltonian.h:226-245).
    d1.assign(ns, 0.0);
    d2.assign((size_t)ns * ns, 0.0);
    for (int i = 0; i < ns; ++i) d1[i] = f.h1[i * ns + i];
    for (int i = 0; i < ns; ++i)
      for (int j = i + 1; j < ns; ++j) d2[i * ns + j] = f.get2(i, j, i, j) - f.get2(i, j, j, i);
  }

  double diagonal(const std::vector<int>& occ) const {
    double v = 0.0;
    for (size_t x = 0; x < occ.size(); ++x) {
      int i = occ[x];
      v += d1[i];
      for (size_t y = x + 1; y < occ.size(); ++y) v += d2[i * ns + occ[y]];
    }
    v += core;
    return v;
  }

  // get_column (hamiltonian.h:412-418): [diagonal, singles, doubles].
  void column(const Det& D, std::vector<std::pair<Det, double>>& out) const {
    out.clear();
    std::vector<int> occ = occupied(D, ns);
    out.push_back({D, diagonal(occ)});
    for (int i : occ) {
      for (const S1& e : sgl[i]) {
        int a = e.a;
        if (D.occ(a)) continue;
        double v;
        if (mixed) {
          // fp32 sum in the same ascending order as the fp64 one.
          float f32 = (float)e.h;
          for (int k : occ) f32 = f32 + (float)e.two[k];
          v = (double)f32;
        } else {
          v = e.h;
          for (int k : occ) v += e.two[k];
        }
        Det nd = D;
        nd.clr(i);
        if (parity(nd, i, a)) v = -v;
        nd.set(a);
        out.push_back({nd, v});
      }
    }
    for (size_t x = 0; x < occ.size(); ++x)
      for (size_t y = x + 1; y < occ.size(); ++y) {
        int i = occ[x], j = occ[y];
        for (const D2& e : dbl[i * ns + j]) {
          if (!(mixed ? e.ok32 : e.ok64)) continue;
          if (D.occ(e.a) || D.occ(e.b)) continue;
          double v = mixed ? (double)e.v32 : e.v;
          Det nd = D;
          nd.clr(i);
          int s1 = parity(nd, i, e.a);
          nd.set(e.a);
          nd.clr(j);
          int s2 = parity(nd, j, e.b);
          nd.set(e.b);
          if (s1 ^ s2) v = -v;
          out.push_back({nd, v});
        }
      }
  }

  // get_off_diagonal (hamiltonian.h:274-328).
  double off_diagonal(const Det& Da, const Det& Db) const {
    Det x;
    x.w[0] = Da.w[0] ^ Db.w[0];
    x.w[1] = Da.w[1] ^ Db.w[1];
    int pc = __builtin_popcountll(x.w[0]) + __builtin_popcountll(x.w[1]);
    Det from, to;
    for (int w = 0; w < 2; ++w) {
      from.w[w] = x.w[w] & Da.w[w];
      to.w[w] = x.w[w] & Db.w[w];
    }
    std::vector<int> io = occupied(from, ns), ao = occupied(to, ns);
    if (pc == 2) {
      int i = io[0];
      for (const S1& e : sgl[i]) {
        if (e.a != ao[0]) continue;
        double v = e.h;
        for (int k = 0; k < ns; ++k)
          if (Da.occ(k)) v += e.two[k];
        Det nd = Da;
        nd.clr(i);
        if (parity(nd, i, e.a)) v = -v;
        return v;
      }
    } else if (pc == 4) {
      int i = io[0], j = io[1];
      for (const D2& e : dbl[i * ns + j]) {
        if (!e.ok64) continue;
        if ((e.a == ao[1] && e.b == ao[0]) || (e.a == ao[0] && e.b == ao[1])) {
          double v = e.v;
          Det nd = Da;
          nd.clr(i);
          int s1 = parity(nd, i, e.a);
          nd.set(e.a);
          nd.clr(j);
          int s2 = parity(nd, j, e.b);
          if (s1 ^ s2) v = -v;
          return v;
        }
      }
    }
    return 0.0;
  }

  // get_hartree_fock (hamiltonian.h:447-492).
  Det hartree_fock() const {
    std::vector<std::pair<int, double>> so;
    if (!has_orb_e) {
      for (int i = 0; i < ns; ++i) so.push_back({i, (double)i});
    } else {
      so = orb_e;
    }
    std::stable_sort(so.begin(), so.end(),
                     [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
                       if (x.second != y.second) return x.second < y.second;
                       return x.first < y.first;
                     });
    Det hf;
    int na = (nelec + ms2) / 2, nb = (nelec - ms2) / 2;
    for (int ia = 0, ib = 0, i = 0; i < nelec && i < (int)so.size(); ++i) {
      int orb = so[i].first;
      if (orb % 2 == 0 && ia < na) {
        hf.set(orb);
        ++ia;
      } else if (orb % 2 == 1 && ib < nb) {
        hf.set(orb);
        ++ib;
      }
    }
    return hf;
  }
};

// ---------------------------------------------------------------- table
struct Entry {
  Det key;
  double x = 0.0;
  i128 z = 0;
  long ins_iter = 0;
  bool used = false;
};

// Open-addressing map Det -> Entry (one shard; single writer).
struct Shard {
  std::vector<Entry> slots;
  size_t count = 0;
  Shard() { slots.resize(64); }
  Entry* find(const Det& d) {
    size_t mask = slots.size() - 1;
    size_t h = det_hash(d) & mask;
    while (slots[h].used) {
      if (slots[h].key == d) return &slots[h];
      h = (h + 1) & mask;
    }
    return nullptr;
  }
  Entry* insert(const Det& d) {
    if ((count + 1) * 2 > slots.size()) grow();
    size_t mask = slots.size() - 1;
    size_t h = det_hash(d) & mask;
    while (slots[h].used) h = (h + 1) & mask;
    slots[h].used = true;
    slots[h].key = d;
    ++count;
    return &slots[h];
  }
  void grow() {
    std::vector<Entry> old;
    old.swap(slots);
    slots.resize(old.size() * 2);
    count = 0;
    for (auto& e : old)
      if (e.used) {
        Entry* n = insert(e.key);
        *n = e;
      }
  }
};

struct Cand {
  Det d;
  double x, z;
};

// ---------------------------------------------------------------- solver
struct Solver {
  const Ham* h = nullptr;
  int k = 1, M = 100;
  int update = 0;  // 0: eig (solver.h:333-609), 1: line search (solver.h:254-314, k = 1)
  double z_thr = 0.0;
  double max_size = 0.0;  // the reference's max_wavefunction_size
  int nthreads = 1;
  std::vector<Shard> shards;
  dd xx{1.0, 0.0}, sc{1.0, 0.0};
  i128 xz = 0;
  long x_size = 0, z_size = 0, updates = 0, cols_size = 0;
  long iter_stamp = 0;
  long iterations = 0;
  std::vector<double> dx_last;
  std::vector<Cand> cands;
  EigStats st;
  int error = 0;  // 0 ok, 1 table full, 2 numeric, 3 too few candidates
  std::string errmsg;
  std::vector<std::vector<std::pair<Det, double>>> cols;
  // phase timers (pick, submatrix, columns, eig, update_x, update_z)
  double timers[6] = {0, 0, 0, 0, 0, 0};

  size_t owner(const Det& d) const { return (det_hash(d) >> 40) % shards.size(); }
  Entry* find(const Det& d) { return shards[owner(d)].find(d); }
  Entry* insert(const Det& d) { return shards[owner(d)].insert(d); }

  dd xss() const { return dd_mul(dd_mul(xx, sc), sc); }
  static double key_of(dd s, double x, double z) {
    dd r = dd_add_d(dd_mul_d(s, x), z);
    return std::fabs(r.hi);
  }
  bool gate(double dz) const { return dd_abs_gt(dd_mul_d(sc, dz), z_thr); }

  void update_x(const std::vector<Det>& picked, const std::vector<double>& dx) {
    for (size_t t = 0; t < picked.size(); ++t) {
      Entry* e = find(picked[t]);
      if (!e) {
        e = insert(picked[t]);
        e->ins_iter = 0;
        e->x = dx[t];
        x_size += 1;
        continue;
      }
      if (!(std::fabs(e->x) != 0.0)) x_size += 1;
      e->x = e->x + dx[t];
      xz += to_fixed(fixed_to_double(e->z) * dx[t]);
    }
  }

  // Canonical update_z_and_get_sub (wavefunction.h:468-605).
  void update_z(const std::vector<Det>& picked, const std::vector<double>& dx) {
    int kk = (int)picked.size();
    ++iter_stamp;
    const long it = iter_stamp;
    dd s = xss();
    size_t P = shards.size();
    std::vector<long> n_new(P, 0), n_upd(P, 0);
    // Scatter: every shard owner walks all elements it owns.
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
    for (size_t p = 0; p < P; ++p) {
      for (int t = 0; t < kk; ++t)
        for (auto& el : cols[t]) {
          if (owner(el.first) != p) continue;
          double dz = el.second * dx[t];
          i128 fz = to_fixed(dz);
          Entry* e = shards[p].find(el.first);
          if (gate(dz)) {
            if (!e) {
              e = shards[p].insert(el.first);
              e->ins_iter = it;
              n_new[p]++;
            }
            e->z += fz;
            n_upd[p]++;
          } else if (e && e->ins_iter != it) {
            e->z += fz;
            n_upd[p]++;
          }
        }
    }
    // Per-column sums and top-M selection on the post-scatter values.
    std::vector<i128> newz(kk, 0), xzt(kk, 0);
    std::vector<std::vector<Cand>> sel(kk);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int t = 0; t < kk; ++t) {
      struct R {
        double val;
        Det d;
        bool present;
      };
      std::vector<R> rk;
      rk.reserve(cols[t].size());
      i128 nz = 0, xt = 0;
      for (auto& el : cols[t]) {
        Entry* e = find(el.first);
        double x = e ? e->x : 0.0;
        double dz = el.second * dx[t];
        nz += to_fixed(el.second * x);
        xt += to_fixed(dz * x);
        bool vis = e && (gate(dz) || e->ins_iter != it);
        double val = vis ? key_of(s, e->x, fixed_to_double(e->z)) : 0.0;
        rk.push_back({val, el.first, e != nullptr});
      }
      auto cmp = [](const R& a, const R& b) {
        if (a.val != b.val) return a.val > b.val;
        return det_greater(a.d, b.d);
      };
      size_t m = std::min((size_t)M, rk.size());
      std::partial_sort(rk.begin(), rk.begin() + m, rk.end(), cmp);
      for (size_t i = 0; i < m; ++i)
        if (rk[i].present) sel[t].push_back({rk[i].d, 0.0, 0.0});
      newz[t] = nz;
      xzt[t] = xt;
    }
    // Reinsert z of the picked determinants, then refetch.
    for (int t = 0; t < kk; ++t) {
      Entry* e = find(picked[t]);
      if (!e) {
        e = insert(picked[t]);
        e->ins_iter = it;
      }
      e->z = newz[t];
    }
    cands.clear();
    for (int t = 0; t < kk; ++t)
      for (auto& c : sel[t]) {
        Entry* e = find(c.d);
        cands.push_back({c.d, e->x, fixed_to_double(e->z)});
      }
    for (int t = 0; t < kk; ++t) xz += xzt[t];
    long nn = 0, nu = 0;
    for (size_t p = 0; p < P; ++p) {
      nn += n_new[p];
      nu += n_upd[p];
    }
    z_size += nn;
    updates = nu;
    if ((double)z_size > 0.79 * max_size) {
      error = 1;
      errmsg = "The hash table is full. Please increase max_wavefunction_size or z_threshold.";
    }
  }

  // gcd_grad pick (solver.h:196-224), canonical tie order.
  bool pick(std::vector<Det>& pd, std::vector<double>& px, std::vector<double>& pz) {
    dd s = xss();
    std::vector<std::pair<double, size_t>> keyed;
    for (size_t i = 0; i < cands.size(); ++i)
      keyed.push_back({key_of(s, cands[i].x, cands[i].z), i});
    std::sort(keyed.begin(), keyed.end(),
              [&](const std::pair<double, size_t>& a, const std::pair<double, size_t>& b) {
                if (a.first != b.first) return a.first > b.first;
                return det_greater(cands[a.second].d, cands[b.second].d);
              });
    pd.clear();
    px.clear();
    pz.clear();
    for (auto& kv : keyed) {
      const Cand& c = cands[kv.second];
      if (!pd.empty() && pd.back() == c.d) continue;
      pd.push_back(c.d);
      px.push_back(c.x);
      pz.push_back(c.z);
      if ((int)pd.size() == k) break;
    }
    return (int)pd.size() == k;
  }

  void init(const Det& ref) {
    shards.assign(std::max(1, nthreads), Shard());
    xx = dd_make(1.0);
    sc = dd_make(1.0);
    xz = 0;
    x_size = z_size = updates = 0;
    iter_stamp = 0;
    iterations = 0;
    error = 0;
    st = EigStats();
    std::vector<Det> picked{ref};
    std::vector<double> dx{1.0};
    update_x(picked, dx);
    cols.assign(1, {});
    h->column(ref, cols[0]);
    update_z(picked, dx);
    cols_size = (long)cols[0].size();
    dx_last = dx;
  }

  int step() {
    if (error) return error;
    double t0 = omp_get_wtime();
    std::vector<Det> pd;
    std::vector<double> px, pz;
    if (!pick(pd, px, pz)) {
      error = 3;
      errmsg = "Fewer than num_of_coordinate distinct candidates are available.";
      return error;
    }
    double t1 = omp_get_wtime();
    std::vector<double> Hs((size_t)k * k, 0.0);
    std::vector<std::vector<int>> occs(k);
#pragma omp parallel for num_threads(nthreads)
    for (int i = 0; i < k; ++i) {
      occs[i] = occupied(pd[i], h->ns);
      for (int j = i + 1; j < k; ++j) Hs[i * k + j] = h->off_diagonal(pd[i], pd[j]);
      Hs[i * k + i] = h->diagonal(occs[i]);
    }
    double t2 = omp_get_wtime();
    cols.assign(k, {});
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int t = 0; t < k; ++t) h->column(pd[t], cols[t]);
    double t3 = omp_get_wtime();
    std::vector<double> dx(k);
    last_pd = pd;
    last_px = px;
    last_pz = pz;
    last_Hs = Hs;
    last_xx = xx;
    last_sc = sc;
    last_xz = xz;
    int rc = EIG_OK;
    if (update == 1) dx[0] = ls_update(Hs[0], px[0], pz[0], xx, sc);
    else rc = eig_update(k, Hs.data(), px.data(), pz.data(), xx, sc, fixed_to_dd(xz), dx.data(), st);
    if (rc != EIG_OK) {
      error = 2;
      errmsg = "The subspace eigenvalue is not negative; sqrt(-lambda) is undefined. The Hamiltonian "
               "(core energy included) must have a negative lowest eigenvalue.";
      return error;
    }
    double t4 = omp_get_wtime();
    update_x(pd, dx);
    double t5 = omp_get_wtime();
    update_z(pd, dx);
    double t6 = omp_get_wtime();
    cols_size = 0;
    for (auto& c : cols) cols_size += (long)c.size();
    dx_last = dx;
    ++iterations;
    timers[0] += t1 - t0;
    timers[1] += t2 - t1;
    timers[2] += t3 - t2;
    timers[3] += t4 - t3;
    timers[4] += t5 - t4;
    timers[5] += t6 - t5;
    return error;
  }

  double energy() const { return dd_div(fixed_to_dd(xz), xx).hi; }

  // inputs of the last eig update (diagnostics)
  std::vector<Det> last_pd;
  std::vector<double> last_px, last_pz, last_Hs;
  dd last_xx{0, 0}, last_sc{0, 0};
  i128 last_xz = 0;
};

struct Ctx {
  Fcidump f;
  Ham h;
  Solver s;
  std::string err;
};

static void det_in(const uint64_t* w, int nw, Det& d) {
  d.w[0] = w[0];
  d.w[1] = nw > 1 ? w[1] : 0;
}
static void det_out(const Det& d, int nw, uint64_t* w) {
  w[0] = d.w[0];
  if (nw > 1) w[1] = d.w[1];
}

}  // namespace oc

using namespace oc;

static std::string g_err;

extern "C" {

const char* oc_last_error() { return g_err.c_str(); }

void* oc_create(const char* fcidump, double threshold) {
  try {
    Ctx* c = new Ctx();
    c->f.load(fcidump);
    if (c->f.norb > 128) throw std::invalid_argument("the oracle supports at most 128 spin-orbitals");
    c->h.build(c->f, threshold);
    c->s.h = &c->h;
    return c;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void oc_destroy(void* p) { delete (Ctx*)p; }

// out: {nwords, spin-orbitals, nelec, ms2}
void oc_info(void* p, int* out) {
  Ctx* c = (Ctx*)p;
  out[0] = c->h.nw;
  out[1] = c->h.ns;
  out[2] = c->h.nelec;
  out[3] = c->h.ms2;
}

double oc_core(void* p) { return ((Ctx*)p)->h.core; }

// Column precision of the following column evaluations: 0 fp64, 1 mixed
// (fp32 off-diagonal column elements, as the device's mixed precision).
void oc_set_precision(void* p, int mixed) { ((Ctx*)p)->h.mixed = mixed != 0; }

void oc_hf(void* p, uint64_t* det) {
  Ctx* c = (Ctx*)p;
  det_out(c->h.hartree_fock(), c->h.nw, det);
}

double oc_diagonal(void* p, const uint64_t* det) {
  Ctx* c = (Ctx*)p;
  Det d;
  det_in(det, c->h.nw, d);
  return c->h.diagonal(occupied(d, c->h.ns));
}

double oc_off_diagonal(void* p, const uint64_t* d1, const uint64_t* d2) {
  Ctx* c = (Ctx*)p;
  Det a, b;
  det_in(d1, c->h.nw, a);
  det_in(d2, c->h.nw, b);
  return c->h.off_diagonal(a, b);
}

long oc_column(void* p, const uint64_t* det, uint64_t* dets, double* vals, long cap) {
  Ctx* c = (Ctx*)p;
  Det d;
  det_in(det, c->h.nw, d);
  std::vector<std::pair<Det, double>> col;
  c->h.column(d, col);
  for (long i = 0; i < (long)col.size() && i < cap; ++i) {
    det_out(col[i].first, c->h.nw, dets + i * c->h.nw);
    vals[i] = col[i].second;
  }
  return (long)col.size();
}

// Canonical lowest eigenpair (for pinning against LAPACK/numpy).
// The subspace solve from e_0 with no previous solve (convergence tested
// from 4 Lanczos vectors on); lambda is the Rayleigh quotient.
void oc_lowest_eig(int n, const double* A, double* lambda, double* q_hi, double* q_lo) {
  std::vector<double> v0(n, 0.0);
  v0[0] = 1.0;
  EigStats st;
  EigSolve S = solve_lowest(n, [&](int r, int c) { return A[(size_t)r * n + c]; }, v0, st, eig_classes(n - 1));
  const EigPair& ep = S.ep;
  *lambda = S.rq.hi;
  for (int i = 0; i < n; ++i) {
    q_hi[i] = ep.q[i].hi;
    q_lo[i] = ep.q[i].lo;
  }
}

// Canonical subspace update.  sc_xx_xz: in {xx.hi, xx.lo, sc.hi, sc.lo, xz.hi, xz.lo};
// out the same layout (xz unchanged).  Returns 0 ok, 1 non-negative eigenvalue.
int oc_eig_update(int k, const double* Hs, const double* x, const double* z, double* s, double* dx,
                  double* stats) {
  dd xx{s[0], s[1]}, sc{s[2], s[3]};
  dd xz = fixed_to_dd(to_fixed(s[4]) + to_fixed(s[5]));
  EigStats st;
  int rc = eig_update(k, Hs, x, z, xx, sc, xz, dx, st);
  s[0] = xx.hi;
  s[1] = xx.lo;
  s[2] = sc.hi;
  s[3] = sc.lo;
  if (stats) {
    stats[0] = st.max_residual;
    stats[1] = st.curr_max_abs_evec;
    stats[2] = st.curr_diff_inv_iter;
  }
  return rc;
}

// Diagnostics of one normal-branch subspace solve on given inputs (fresh
// schedule): out = {first attempt's Lanczos vectors, attempts, final
// attempt's vectors, alpha_0, theta, |T|_inf, residual}.
void oc_eig_trace(int k, const double* Hs, const double* x, const double* z, const double* s, double* out) {
  auto H = [&](int r, int c) { return r <= c ? Hs[r * k + c] : Hs[c * k + r]; };
  const int cls = eig_classes(k);
  std::vector<double> Hx(k);
  for (int r = 0; r < k; ++r) Hx[r] = dotc(cls, k, [&](int c) { return H(r, c); }, [&](int c) { return x[c]; });
  double xTx, xTz, xTHx;
  setup_sums(cls, k, x, z, Hx.data(), xTx, xTz, xTHx);
  dd xx{s[0], s[1]};
  dd xz = fixed_to_dd(to_fixed(s[4]) + to_fixed(s[5]));
  const double yy = dd_add_d(xx, -xTx).hi, y1 = std::sqrt(yy);
  const int n = k + 1;
  const double a00 = dd_div_d(dd_add_d(dd_add_d(xz, -(2.0 * xTz)), xTHx), yy).hi;
  std::vector<double> b(k), v0(n);
  for (int i = 0; i < k; ++i) b[i] = (z[i] - Hx[i]) / y1;
  auto A = [&](int r, int c) -> double {
    if (r == 0) return c == 0 ? a00 : b[c - 1];
    if (c == 0) return b[r - 1];
    return H(r - 1, c - 1);
  };
  v0[0] = y1;
  for (int i = 0; i < k; ++i) v0[i + 1] = x[i];
  EigStats st;
  EigSolve S = solve_lowest(n, A, v0, st, cls);
  out[0] = S.steps_first;
  out[1] = S.attempts;
  out[2] = S.ep.steps;
  out[3] = S.ep.alpha0;
  out[4] = S.ep.lambda;
  out[5] = S.ep.tn;
  out[6] = std::sqrt(S.res2);
  out[7] = S.Ae[0];
  out[8] = S.ev[0];
  out[9] = S.rq.hi;
  out[10] = S.Ae[n - 1];
  out[11] = S.ev[n - 1];
}

// Solver. ref_det may be NULL (Hartree-Fock).
int oc_solver_init(void* p, int k, int M, double z_thr, double max_size, int nthreads,
                   const uint64_t* ref_det) {
  Ctx* c = (Ctx*)p;
  Solver& s = c->s;
  s.k = k;
  s.M = M;
  s.z_thr = z_thr;
  s.max_size = max_size;
  s.nthreads = nthreads > 0 ? nthreads : 1;
  Det ref;
  if (ref_det) det_in(ref_det, c->h.nw, ref);
  else ref = c->h.hartree_fock();
  s.init(ref);
  return s.error;
}

// Coordinate update of the following steps: 0 eig, 1 line search (k = 1).
int oc_solver_set_update(void* p, int mode) {
  Ctx* c = (Ctx*)p;
  if (mode == 1 && c->s.k != 1) {
    g_err = "the line-search update takes one coordinate per iteration";
    return 1;
  }
  c->s.update = mode;
  return 0;
}

// The restated line-search step on given inputs: out[0] = dx, out[1..2] =
// the new xx (hi, lo).
void oc_ls_update(double h, double x, double z, double xx_hi, double xx_lo, double* out) {
  dd xx{xx_hi, xx_lo}, sc = dd_make(1.0);
  out[0] = ls_update(h, x, z, xx, sc);
  out[1] = xx.hi;
  out[2] = xx.lo;
}

int oc_solver_step(void* p, long n) {
  Ctx* c = (Ctx*)p;
  for (long i = 0; i < n; ++i) {
    int rc = c->s.step();
    if (rc) {
      g_err = c->s.errmsg;
      return rc;
    }
  }
  return 0;
}

// out[0..16]: energy, dx_max, x_size, z_size, cols_size, updates, scale.hi,
// xx.hi, xx.lo, scale.lo, xz (double), iterations, max_residual,
// curr_max_abs_evec, curr_diff_inv_iter, error, unconverged subspace solves
void oc_solver_report(void* p, double* out) {
  Ctx* c = (Ctx*)p;
  Solver& s = c->s;
  double dxm = 0.0;
  for (double v : s.dx_last) dxm = std::max(dxm, std::fabs(v));
  out[0] = s.energy();
  out[1] = dxm;
  out[2] = (double)s.x_size;
  out[3] = (double)s.z_size;
  out[4] = (double)s.cols_size;
  out[5] = (double)s.updates;
  out[6] = s.sc.hi;
  out[7] = s.xx.hi;
  out[8] = s.xx.lo;
  out[9] = s.sc.lo;
  out[10] = fixed_to_double(s.xz);
  out[11] = (double)s.iterations;
  out[12] = s.st.max_residual;
  out[13] = s.st.curr_max_abs_evec;
  out[14] = s.st.curr_diff_inv_iter;
  out[15] = (double)s.error;
  out[16] = (double)s.st.eig_unconverged;
}

void oc_solver_timers(void* p, double* out) {
  Ctx* c = (Ctx*)p;
  for (int i = 0; i < 6; ++i) out[i] = c->s.timers[i];
}

// Table dump sorted by determinant (ascending).  z as double; zfix as
// {hi64, lo64} of the fixed-point value when zfix != NULL.
long oc_table_dump(void* p, uint64_t* dets, double* x, double* z, int64_t* zfix, long cap) {
  Ctx* c = (Ctx*)p;
  std::vector<const Entry*> all;
  for (auto& sh : c->s.shards)
    for (auto& e : sh.slots)
      if (e.used) all.push_back(&e);
  std::sort(all.begin(), all.end(), [](const Entry* a, const Entry* b) {
    if (a->key.w[1] != b->key.w[1]) return a->key.w[1] < b->key.w[1];
    return a->key.w[0] < b->key.w[0];
  });
  int nw = c->h.nw;
  for (long i = 0; i < (long)all.size() && i < cap; ++i) {
    det_out(all[i]->key, nw, dets + i * nw);
    x[i] = all[i]->x;
    z[i] = fixed_to_double(all[i]->z);
    if (zfix) {
      zfix[2 * i] = (int64_t)(all[i]->z >> 64);
      zfix[2 * i + 1] = (int64_t)(uint64_t)all[i]->z;
    }
  }
  return (long)all.size();
}

// Order-independent digest of the table, the definition of
// mcdfci_table_digest: {entries, sum of key hashes, sum of entry hashes}.
void oc_table_digest(void* p, uint64_t* out) {
  Ctx* c = (Ctx*)p;
  uint64_t cnt = 0, ks = 0, es = 0;
  for (auto& sh : c->s.shards)
    for (auto& en : sh.slots) {
      if (!en.used) continue;
      uint64_t e = mix64(en.key.w[0] + 0x9e3779b97f4a7c15ULL);
      e = mix64(e ^ en.key.w[1]);
      ks += e;
      uint64_t xb = 0;
      if (en.x != 0.0) std::memcpy(&xb, &en.x, 8);
      e = mix64(e ^ xb);
      e = mix64(e ^ (uint64_t)(en.z >> 64));
      es += mix64(e ^ (uint64_t)en.z);
      ++cnt;
    }
  out[0] = cnt;
  out[1] = ks;
  out[2] = es;
}

long oc_candidates(void* p, uint64_t* dets, double* x, double* z, long cap) {
  Ctx* c = (Ctx*)p;
  int nw = c->h.nw;
  auto& cs = c->s.cands;
  for (long i = 0; i < (long)cs.size() && i < cap; ++i) {
    det_out(cs[i].d, nw, dets + i * nw);
    x[i] = cs[i].x;
    z[i] = cs[i].z;
  }
  return (long)cs.size();
}

// Inputs and output of the last iteration's eig update: dets[k], x[k],
// z[k], Hs[k*k], s6 = {xx.hi, xx.lo, sc.hi, sc.lo, xz_fixed_hi, xz_fixed_lo}, dx[k].
int oc_last_step(void* p, uint64_t* dets, double* x, double* z, double* Hs, double* s6, double* dx) {
  Ctx* c = (Ctx*)p;
  Solver& s = c->s;
  int k = (int)s.last_pd.size();
  int nw = c->h.nw;
  for (int i = 0; i < k; ++i) {
    det_out(s.last_pd[i], nw, dets + i * nw);
    x[i] = s.last_px[i];
    z[i] = s.last_pz[i];
    dx[i] = s.dx_last[i];
  }
  for (int i = 0; i < k * k; ++i) Hs[i] = s.last_Hs[i];
  s6[0] = s.last_xx.hi;
  s6[1] = s.last_xx.lo;
  s6[2] = s.last_sc.hi;
  s6[3] = s.last_sc.lo;
  int64_t hi = (int64_t)(s.last_xz >> 64);
  uint64_t lo = (uint64_t)s.last_xz;
  std::memcpy(&s6[4], &hi, 8);
  std::memcpy(&s6[5], &lo, 8);
  return k;
}

// Wavefunction-level hooks (update_x / update_z on caller-chosen picks).
void oc_wf_reset(void* p, int k, int M, double z_thr, double max_size, int nthreads) {
  Ctx* c = (Ctx*)p;
  Solver& s = c->s;
  s.k = k;
  s.M = M;
  s.z_thr = z_thr;
  s.max_size = max_size;
  s.nthreads = nthreads > 0 ? nthreads : 1;
  s.shards.assign(s.nthreads, Shard());
  s.xx = dd_make(1.0);
  s.sc = dd_make(1.0);
  s.xz = 0;
  s.x_size = s.z_size = s.updates = 0;
  s.iter_stamp = 0;
  s.error = 0;
  s.cands.clear();
}

// s: {xx.hi, xx.lo, sc.hi, sc.lo}; xz set from a double.
void oc_wf_set_scalars(void* p, const double* v, double xz) {
  Ctx* c = (Ctx*)p;
  c->s.xx = dd{v[0], v[1]};
  c->s.sc = dd{v[2], v[3]};
  c->s.xz = to_fixed(xz);
}

// out: {x_size, z_size, updates, xz, error}
void oc_wf_state(void* p, double* out) {
  Ctx* c = (Ctx*)p;
  out[0] = (double)c->s.x_size;
  out[1] = (double)c->s.z_size;
  out[2] = (double)c->s.updates;
  out[3] = fixed_to_double(c->s.xz);
  out[4] = (double)c->s.error;
}

void oc_wf_update(void* p, int k, const uint64_t* dets, const double* dx) {
  Ctx* c = (Ctx*)p;
  int nw = c->h.nw;
  std::vector<Det> pd(k);
  for (int t = 0; t < k; ++t) det_in(dets + t * nw, nw, pd[t]);
  std::vector<double> d(dx, dx + k);
  c->s.update_x(pd, d);
  c->s.cols.assign(k, {});
  for (int t = 0; t < k; ++t) c->h.column(pd[t], c->s.cols[t]);
  c->s.update_z(pd, d);
}

// Canonical pick on a caller-provided candidate list; returns the number picked.
int oc_pick(void* p, long n, const uint64_t* dets, const double* x, const double* z, const double* s4, int k,
            uint64_t* out) {
  Ctx* c = (Ctx*)p;
  int nw = c->h.nw;
  Solver& s = c->s;
  s.cands.clear();
  for (long i = 0; i < n; ++i) {
    Cand cd;
    det_in(dets + i * nw, nw, cd.d);
    cd.x = x[i];
    cd.z = z[i];
    s.cands.push_back(cd);
  }
  s.xx = dd{s4[0], s4[1]};
  s.sc = dd{s4[2], s4[3]};
  int kk = s.k;
  s.k = k;
  std::vector<Det> pd;
  std::vector<double> px, pz;
  s.pick(pd, px, pz);
  s.k = kk;
  for (size_t i = 0; i < pd.size(); ++i) det_out(pd[i], nw, out + i * nw);
  return (int)pd.size();
}

// Numeric helpers exposed for pinning tests.
void oc_to_fixed(double d, int64_t* out) {
  i128 v = to_fixed(d);
  out[0] = (int64_t)(v >> 64);
  out[1] = (int64_t)(uint64_t)v;
}
double oc_fixed_to_double(int64_t hi, int64_t lo) {
  i128 v = ((i128)hi << 64) | (i128)(uint64_t)lo;
  return fixed_to_double(v);
}
void oc_dd_ops(double ah, double al, double bh, double bl, double* out) {
  dd a{ah, al}, b{bh, bl};
  dd r;
  r = dd_add(a, b);
  out[0] = r.hi;
  out[1] = r.lo;
  r = dd_mul(a, b);
  out[2] = r.hi;
  out[3] = r.lo;
  r = dd_div(a, b);
  out[4] = r.hi;
  out[5] = r.lo;
  r = dd_sqrt(a);
  out[6] = r.hi;
  out[7] = r.lo;
}

// Gradient of the variational energy of the wavefunction (dets, x) with
// respect to the unique integrals (canonical positions; see
// include/mcdfci.h, mcdfci_energy_gradient).  An independent restatement:
// excitations are enumerated over spin-orbitals (every single and double of
// each determinant, no threshold), the matrix-element terms are the
// Slater-Condon expressions of src/hamiltonian.h:257-410 written with
// spatial chemists' integrals, and all sums are exact fixed-point integers.
// Returns 0, or 1 when x is zero.
int oc_energy_gradient(void* p, const uint64_t* dets, const double* x, long n, double* grad_h,
                       double* grad_eri, double* energy) {
  Ctx* c = (Ctx*)p;
  const Fcidump& f = c->f;
  const int ns = f.norb, M = ns / 2, nw = c->h.nw;
  const size_t nh = (size_t)M * M, n4 = nh * nh;
  std::vector<Det> D(n);
  for (long i = 0; i < n; ++i) det_in(dets + (size_t)i * nw, nw, D[i]);
  std::unordered_map<uint64_t, std::vector<std::pair<Det, double>>> xm;
  for (long i = 0; i < n; ++i) xm[det_hash(D[i])].push_back({D[i], x[i]});
  auto lookup = [&](const Det& d) -> double {
    auto it = xm.find(det_hash(d));
    if (it == xm.end()) return 0.0;
    for (auto& e : it->second)
      if (e.first == d) return e.second;
    return 0.0;
  };
  i128 s1 = 0;
  for (long i = 0; i < n; ++i) s1 += to_fixed(x[i] * x[i]);
  const double xx = fixed_to_double(s1);
  if (!(xx > 0.0)) return 1;
  const double scale = 1.0 / std::sqrt(xx);
  i128 q = 0;
  for (long i = 0; i < n; ++i) {
    const double ci = x[i] * scale;
    q += to_fixed(ci * ci);
  }
  // Exact fixed-point sums, so the per-thread partial arrays can be added
  // in any order.
  std::vector<i128> gh_all(nh, 0), ge_all(n4, 0);
  auto hi = [&](int a, int b) { return a >= b ? (size_t)a * M + b : (size_t)b * M + a; };
  auto ei = [&](int a, int b, int cc, int d) {
    int p1 = std::max(a, b), q1 = std::min(a, b), p2 = std::max(cc, d), q2 = std::min(cc, d);
    if (p1 * (p1 + 1) / 2 + q1 < p2 * (p2 + 1) / 2 + q2) {
      std::swap(p1, p2);
      std::swap(q1, q2);
    }
    return (((size_t)p1 * M + q1) * M + p2) * M + q2;
  };
#pragma omp parallel
  {
  std::vector<i128> gh_t(nh, 0), ge_t(n4, 0);
  auto& gh = gh_t;
  auto& ge = ge_t;
#pragma omp for schedule(dynamic, 256)
  for (long di = 0; di < n; ++di) {
    const Det& d = D[di];
    const double cD = x[di] * scale;
    const std::vector<int> occ = occupied(d, ns);
    std::vector<int> vir;
    for (int s = 0; s < ns; ++s)
      if (!d.occ(s)) vir.push_back(s);
    {
      const i128 w = to_fixed(cD * cD);
      for (size_t a = 0; a < occ.size(); ++a) {
        const int I = occ[a];
        gh[hi(I >> 1, I >> 1)] += w;
        for (size_t b = a + 1; b < occ.size(); ++b) {
          const int J = occ[b];
          ge[ei(I >> 1, I >> 1, J >> 1, J >> 1)] += w;
          if ((I & 1) == (J & 1)) ge[ei(I >> 1, J >> 1, J >> 1, I >> 1)] -= w;
        }
      }
    }
    for (int I : occ)
      for (int A : vir) {
        if ((I & 1) != (A & 1)) continue;
        Det e = d;
        e.clr(I);
        const int sg = parity(e, I, A);
        e.set(A);
        const double xo = lookup(e);
        if (!(std::fabs(xo) != 0.0)) continue;
        i128 w = to_fixed(cD * (xo * scale));
        if (sg) w = -w;
        const int pI = I >> 1, pA = A >> 1;
        gh[hi(pI, pA)] += w;
        for (int K : occ) {
          ge[ei(pI, pA, K >> 1, K >> 1)] += w;
          if ((K & 1) == (I & 1)) ge[ei(pI, K >> 1, K >> 1, pA)] -= w;
        }
      }
    for (size_t a = 0; a < occ.size(); ++a)
      for (size_t b = a + 1; b < occ.size(); ++b) {
        const int I = occ[a], J = occ[b];
        for (size_t u = 0; u < vir.size(); ++u)
          for (size_t v = u + 1; v < vir.size(); ++v) {
            int A = vir[u], B = vir[v];
            const int sI = I & 1, sJ = J & 1;
            if (sI == sJ) {
              if ((A & 1) != sI || (B & 1) != sI) continue;
            } else {
              if ((A & 1) == (B & 1)) continue;
              if ((A & 1) != sI) std::swap(A, B);  // A pairs with I by spin
            }
            Det e = d;
            e.clr(I);
            int sg = parity(e, I, A);
            e.set(A);
            e.clr(J);
            sg ^= parity(e, J, B);
            e.set(B);
            const double xo = lookup(e);
            if (!(std::fabs(xo) != 0.0)) continue;
            i128 w = to_fixed(cD * (xo * scale));
            if (sg) w = -w;
            ge[ei(I >> 1, A >> 1, J >> 1, B >> 1)] += w;
            if (sI == sJ) ge[ei(I >> 1, B >> 1, J >> 1, A >> 1)] -= w;
          }
      }
  }
#pragma omp critical
  {
    for (size_t i = 0; i < nh; ++i) gh_all[i] += gh_t[i];
    for (size_t i = 0; i < n4; ++i) ge_all[i] += ge_t[i];
  }
  }
  const dd den = fixed_to_dd(q);
  double e = c->h.core;
  const std::vector<i128>& gh = gh_all;
  const std::vector<i128>& ge = ge_all;
  for (int a = 0; a < M; ++a)
    for (int b = 0; b < M; ++b) {
      const size_t i = (size_t)a * M + b;
      grad_h[i] = dd_div(fixed_to_dd(gh[i]), den).hi;
      e += grad_h[i] * f.h1[(size_t)(2 * a) * ns + 2 * b];
    }
  for (size_t i = 0; i < n4; ++i) {
    grad_eri[i] = dd_div(fixed_to_dd(ge[i]), den).hi;
    if (grad_eri[i] != 0.0) {
      const int a = (int)(i / (nh * M)), b = (int)(i / nh % M), cc = (int)(i / M % M), d = (int)(i % M);
      // (ab|cd) = <ac|bd>; the store keeps <ij|..> with i <= j (<ij|ab> = <ji|ba>)
      const double v = a <= cc ? f.get2(2 * a, 2 * cc, 2 * b, 2 * d) : f.get2(2 * cc, 2 * a, 2 * d, 2 * b);
      e += grad_eri[i] * v;
    }
  }
  if (energy) *energy = e;
  return 0;
}

}  // extern "C"
