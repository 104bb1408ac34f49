// FCIDUMP reader for the drop-in surface (same dialect as src/input.h:93-330)
// and the dense spatial integral arrays the device tables are built from.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mc {

struct SpatialHam {
  int norb = 0;   // spatial orbitals
  int nelec = 0;
  int ms2 = 0;
  bool uhf = false;
  double core = 0.0;
  std::vector<double> h;    // norb^2, h[p*norb+q]
  std::vector<double> eri;  // norb^4 chemists' (pq|rs), index ((p*n+q)*n+r)*n+s
  // Orbital energies per spin-orbital (interleaved), first insert wins
  // (src/input.h:273-277); empty when the file carries none.
  std::vector<std::pair<int, double>> orb_e;
};

// Throws std::invalid_argument with the reference's messages on bad input.
SpatialHam read_fcidump(const std::string& path, bool* warned_uhf_default);

// Hartree-Fock determinant rule of src/hamiltonian.h:447-492 (interleaved
// bit string, up to 2 words).
void hartree_fock(const SpatialHam& h, uint64_t det[2], bool* warned_no_orbital_energy);

}  // namespace mc
