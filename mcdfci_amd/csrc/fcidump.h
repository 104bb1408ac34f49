// FCIDUMP reader for the drop-in surface (same dialect as src/input.h:93-330)
// and the dense spatial integral arrays the device tables are built from.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mc {

// Integrals by spin block.  Spin-orbital 2p + s is orbital p of spin s
// (0 alpha, 1 beta), the reference's interleaving (src/input.h:242-243).
// Spin-restricted (uhf false): h and eri hold every block.  Unrestricted
// (an FCIDUMP with UHF=.TRUE., whose indices are spin-orbitals,
// src/input.h:125-127,203-237): h / hb are the alpha / beta one-electron
// integrals and eri / eri_bb / eri_ab the (p_a q_a|r_a s_a),
// (p_b q_b|r_b s_b) and (p_a q_a|r_b s_b) blocks; the (bb|aa) block is
// eri_ab with the pairs exchanged.
struct SpatialHam {
  int norb = 0;   // spatial orbitals (half the spin-orbitals)
  int nelec = 0;
  int ms2 = 0;
  bool uhf = false;
  double core = 0.0;
  std::vector<double> h;    // norb^2, h[p*norb+q] (alpha block when uhf)
  std::vector<double> eri;  // norb^4 chemists' (pq|rs), index ((p*n+q)*n+r)*n+s (alpha-alpha when uhf)
  std::vector<double> hb, eri_bb, eri_ab;  // uhf only
  // Orbital energies per spin-orbital (interleaved), first insert wins
  // (src/input.h:230-233,273-277); empty when the file carries none.
  std::vector<std::pair<int, double>> orb_e;

  const double* h_spin(int s) const { return uhf && s ? hb.data() : h.data(); }
  // block 0: alpha-alpha, 1: beta-beta, 2: alpha-beta
  const double* eri_block(int b) const { return !uhf || b == 0 ? eri.data() : (b == 1 ? eri_bb.data() : eri_ab.data()); }
};

// Throws std::invalid_argument with the reference's messages on bad input.
SpatialHam read_fcidump(const std::string& path, bool* warned_uhf_default);

// Hartree-Fock determinant rule of src/hamiltonian.h:447-492 (interleaved
// bit string, up to 2 words).
void hartree_fock(const SpatialHam& h, uint64_t det[2], bool* warned_no_orbital_energy);

}  // namespace mc
