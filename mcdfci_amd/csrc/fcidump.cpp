// FCIDUMP reader.  Dialect (src/input.h:93-285): the header is every line
// up to the first one containing "&END" or "/" (that line is dropped);
// NORB, NELEC, MS2 are required integers, UHF is optional; the body is
// "value i a j b" with 1-based indices: i,a,j,b != 0 -> (ia|jb); j = b = 0
// -> h_ia; a = j = b = 0 -> orbital energy of i; all zero -> core energy.
// Two-electron integrals keep the first value seen for each index
// permutation (the reference's unordered_map::insert), one-electron
// integrals and the core energy keep the last.  With UHF=.TRUE. NORB and
// the indices count spin-orbitals (2p + 1 alpha and 2p + 2 beta of orbital
// p, 1-based; src/input.h:125-127,203-237); the integrals are sorted into
// spin blocks, and one that changes the spin of an electron is refused.

#include "fcidump.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <regex>
#include <sstream>
#include <stdexcept>

namespace mc {

namespace {

int header_int(const std::string& header, const std::string& name) {
  std::regex r(name + R"([ ]*=[ ]*(\d+))");
  std::smatch m;
  if (!std::regex_search(header, m, r)) throw std::invalid_argument(name + " is not found.");
  return std::atoi(std::string(m[1]).c_str());
}

}  // namespace

SpatialHam read_fcidump(const std::string& path, bool* warned_uhf_default) {
  SpatialHam H;
  if (warned_uhf_default) *warned_uhf_default = false;
  try {
    std::ifstream f(path);
    if (f.fail()) throw std::invalid_argument("Failed to open the FCIDUMP file " + path);
    std::string header, line;
    bool found_end = false;
    while (std::getline(f, line)) {
      if (line.find("&END") != std::string::npos || line.find("/") != std::string::npos) {
        found_end = true;
        break;
      }
      header += " ";
      header += line;
    }
    header += " &END";
    if (!found_end) throw std::invalid_argument("FCIDUMP has the wrong header.");
    H.norb = header_int(header, "NORB");
    H.nelec = header_int(header, "NELEC");
    H.ms2 = header_int(header, "MS2");
    {
      std::regex r(R"(UHF[ ]*=[ .]*(FALSE|TRUE))");
      std::smatch m;
      if (std::regex_search(header, m, r)) {
        H.uhf = (m[1] == "TRUE");
      } else {
        H.uhf = false;
        if (warned_uhf_default) *warned_uhf_default = true;
      }
    }
    if (H.norb <= 0) throw std::invalid_argument("NORB must be positive.");
    if (H.uhf) {
      // NORB counts spin-orbitals in an unrestricted file (src/input.h:125-127).
      if (H.norb % 2)
        throw std::invalid_argument("an unrestricted FCIDUMP with an odd number of spin-orbitals (NORB) is not "
                                    "supported by the device path.");
      H.norb /= 2;
    }
    const int n = H.norb;      // spatial orbitals
    const int nf = H.uhf ? 2 * n : n;  // largest index in the file
    const size_t n2 = (size_t)n * n, n4 = n2 * n2;
    H.h.assign(n2, 0.0);
    H.eri.assign(n4, 0.0);
    std::vector<uint8_t> seen(n4, 0), seen_bb, seen_ab;
    if (H.uhf) {
      H.hb.assign(n2, 0.0);
      H.eri_bb.assign(n4, 0.0);
      H.eri_ab.assign(n4, 0.0);
      seen_bb.assign(n4, 0);
      seen_ab.assign(n4, 0);
    }
    // Body: read the remainder as one buffer and tokenize with strtod.
    std::stringstream ss;
    ss << f.rdbuf();
    std::string body = ss.str();
    const char* c = body.c_str();
    char* endp = nullptr;
    auto idx = [n](int p, int q, int r, int s) {
      return (((size_t)p * n + q) * n + r) * n + s;
    };
    // One (pq|rs) of spin-orbitals into its spin block; first insert wins.
    // A spin-changing pair (p, q or r, s of different spins) has no block:
    // zero values are dropped, others refused.
    auto put_uhf = [&](int p, int q, int r, int s, double v) {
      if ((p & 1) != (q & 1) || (r & 1) != (s & 1)) {
        if (v != 0.0)
          throw std::invalid_argument("spin-changing two-electron integrals are not supported by the device path.");
        return;
      }
      const int sp = p & 1, sr = r & 1;
      size_t x;
      std::vector<double>* dst;
      std::vector<uint8_t>* sn;
      if (sp == sr) {
        x = idx(p >> 1, q >> 1, r >> 1, s >> 1);
        dst = sp ? &H.eri_bb : &H.eri;
        sn = sp ? &seen_bb : &seen;
      } else {
        x = sp == 0 ? idx(p >> 1, q >> 1, r >> 1, s >> 1) : idx(r >> 1, s >> 1, p >> 1, q >> 1);
        dst = &H.eri_ab;
        sn = &seen_ab;
      }
      if (!(*sn)[x]) {
        (*sn)[x] = 1;
        (*dst)[x] = v;
      }
    };
    while (true) {
      double v = std::strtod(c, &endp);
      if (endp == c) break;
      c = endp;
      long t[4];
      bool ok = true;
      for (int u = 0; u < 4; ++u) {
        t[u] = std::strtol(c, &endp, 10);
        if (endp == c) {
          ok = false;
          break;
        }
        c = endp;
      }
      if (!ok) break;
      long i = t[0], a = t[1], j = t[2], b = t[3];
      if (i < 0 || a < 0 || j < 0 || b < 0 || i > nf || a > nf || j > nf || b > nf)
        throw std::invalid_argument("integral index out of range.");
      if (i && a && j && b) {
        int p = (int)i - 1, q = (int)a - 1, r = (int)j - 1, s = (int)b - 1;
        if (H.uhf) {
          const int perm[8][4] = {{p, q, r, s}, {q, p, r, s}, {p, q, s, r}, {q, p, s, r},
                                  {r, s, p, q}, {s, r, p, q}, {r, s, q, p}, {s, r, q, p}};
          for (auto& e : perm) put_uhf(e[0], e[1], e[2], e[3], v);
          continue;
        }
        const size_t perm[8] = {idx(p, q, r, s), idx(q, p, r, s), idx(p, q, s, r), idx(q, p, s, r),
                                idx(r, s, p, q), idx(s, r, p, q), idx(r, s, q, p), idx(s, r, q, p)};
        for (size_t x : perm)
          if (!seen[x]) {
            seen[x] = 1;
            H.eri[x] = v;
          }
      } else if (a) {
        if (H.uhf) {
          const int p = (int)i - 1, q = (int)a - 1;
          if ((p & 1) != (q & 1)) {
            if (v != 0.0)
              throw std::invalid_argument("spin-changing one-electron integrals are not supported by the device "
                                          "path.");
            continue;
          }
          std::vector<double>& hs = (p & 1) ? H.hb : H.h;
          hs[(size_t)(p >> 1) * n + (q >> 1)] = v;
          hs[(size_t)(q >> 1) * n + (p >> 1)] = v;
          continue;
        }
        H.h[(size_t)(i - 1) * n + (a - 1)] = v;
        H.h[(size_t)(a - 1) * n + (i - 1)] = v;
      } else if (i) {
        // orbital energy of spatial orbital i (both spins), or of
        // spin-orbital i in an unrestricted file
        for (int spin = 0; spin < (H.uhf ? 1 : 2); ++spin) {
          int so = H.uhf ? (int)i - 1 : 2 * ((int)i - 1) + spin;
          bool have = false;
          for (auto& e : H.orb_e)
            if (e.first == so) have = true;
          if (!have) H.orb_e.push_back({so, v});
        }
      } else {
        H.core = v;
      }
    }
  } catch (const std::exception& e) {
    throw std::invalid_argument("Can not read the FCIDUMP file " + path + ". " + e.what());
  }
  return H;
}

void hartree_fock(const SpatialHam& h, uint64_t det[2], bool* warned) {
  det[0] = det[1] = 0;
  std::vector<std::pair<int, double>> so;
  const int ns = 2 * h.norb;
  if (h.orb_e.empty()) {
    for (int i = 0; i < ns; ++i) so.push_back({i, (double)i});
    if (warned) *warned = true;
  } else {
    so = h.orb_e;
    if (warned) *warned = false;
  }
  std::stable_sort(so.begin(), so.end(), [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
    if (x.second != y.second) return x.second < y.second;
    return x.first < y.first;
  });
  int na = (h.nelec + h.ms2) / 2, nb = (h.nelec - h.ms2) / 2;
  for (int ia = 0, ib = 0, i = 0; i < h.nelec && i < (int)so.size(); ++i) {
    int orb = so[i].first;
    if (orb % 2 == 0 && ia < na) {
      det[orb >> 6] |= 1ULL << (orb & 63);
      ++ia;
    } else if (orb % 2 == 1 && ib < nb) {
      det[orb >> 6] |= 1ULL << (orb & 63);
      ++ib;
    }
  }
}

}  // namespace mc
