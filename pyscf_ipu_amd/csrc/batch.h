// Per-molecule descriptors of the batched form: many molecules of uneven size
// share one launch per SCF phase; each kernel finds its molecule's data through
// these offsets into the concatenated device buffers.
#pragma once
#include <stdint.h>

namespace pscf {

// 16 int64 fields, row-major [M][16] on the device (built by the host).
struct BatchMol {
  int64_t natm;    // atoms
  int64_t nbas;    // shells
  int64_t nao;     // AOs (N)
  int64_t nocc;    // doubly occupied orbitals
  int64_t atm0;    // first atom in the global atm array
  int64_t bas0;    // first shell in the global bas array
  int64_t ns, np;  // s and p shell counts
  int64_t shl0;    // offset into the class-sorted shell list (s shells, then p shells)
  int64_t ngrid;   // grid points
  int64_t grid0;   // first grid point in the global grid arrays
  int64_t eri0;    // offset of the packed s8 ERIs (elements)
  int64_t mat0;    // offset of this molecule's N x N matrices (elements)
  int64_t vec0;    // offset of this molecule's length-N vectors
  int64_t ao0;     // offset of this molecule's AO block (4 x N x ngrid elements)
  int64_t state0;  // offset of this molecule's SCF state (doubles)
};

// Per-molecule SCF control, doubles [M][8]:
// [0] iteration, [1] active, [2] converged, [3] E_prev, [4] E_last, [5] gap,
// [6] iterations used, [7] spare.
constexpr int CTL = 8;
// iteration energies kept per molecule by pscf_batch_energy (the last ones)
constexpr int BATCH_HIST = 5;

}  // namespace pscf
