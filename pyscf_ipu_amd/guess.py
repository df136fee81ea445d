"""Initial density guesses (host set-up).

The reference starts every SCF from PySCF's MINAO guess
(`pyscf.scf.hf.init_guess_by_minao`, pyscf_ipu/nanoDFT/nanoDFT.py:233; local
copy pyscf_ipu/pyscf_utils/minao.py:5-123): a block-diagonal superposition of
atomic densities with fractional shell occupations (`atom_hf.frac_occ`) in the
ANO minimal basis, projected onto the molecular basis.  The ANO basis data is
not available here, so `minao` keeps the occupation rule and replaces the ANO
functions by the molecule's own atom-centred shells, orthonormalised within each
atom and angular momentum (Gram-Schmidt in basis order).  The projection step
is then the identity.  Parity with PySCF's MINAO density is unpinned.  For
most molecules the guess changes only the SCF iteration count; a molecule with
a small HOMO-LUMO gap can have several SCF solutions (measured: 1-31 mHa apart),
and there the start decides which one is reached, so which solution such a
molecule reaches from PySCF's MINAO start is not pinned either.
"""
from __future__ import annotations

import numpy as np

from .gto import ANG_OF, ATOM_OF

# Non-relativistic spherical-average occupations per angular momentum (s, p),
# as PySCF's NRSRHF_CONFIGURATION, H..Ne.
_CONFIG = {1: (1, 0), 2: (2, 0), 3: (3, 0), 4: (4, 0), 5: (4, 1), 6: (4, 2), 7: (4, 3), 8: (4, 4), 9: (4, 5),
           10: (4, 6)}


def frac_occ(z: int, l: int):
    """(number of doubly occupied shells, fractional occupation of the next)."""
    ne = _CONFIG[z][l]
    if ne == 0:
        return 0, 0.0
    nd = (2 * l + 1) * 2
    ndocc = ne // nd
    return ndocc, (ne / nd - ndocc) * 2


def minao(mol, S: np.ndarray) -> np.ndarray:
    """Block-diagonal atomic-density guess; `S` is the AO overlap matrix."""
    N = mol.nao_nr()
    ao_loc = mol.ao_loc_nr()
    dm = np.zeros((N, N))
    z = mol.atom_charges()
    for ia in range(mol.natm):
        for l in (0, 1):
            shells = [ib for ib in range(mol.nbas) if mol._bas[ib, ATOM_OF] == ia and mol._bas[ib, ANG_OF] == l]
            if not shells:
                continue
            ndocc, frac = frac_occ(int(z[ia]), l)
            occ = [2.0] * ndocc + [frac]
            ncomp = 2 * l + 1
            for m in range(ncomp):
                idx = np.array([ao_loc[ib] + m for ib in shells])
                Sl = S[np.ix_(idx, idx)]
                # Gram-Schmidt in basis order: columns of C are orthonormal orbitals.
                L = np.linalg.cholesky(Sl)
                C = np.linalg.inv(L).T
                for k, o in enumerate(occ[:len(idx)]):
                    if o:
                        dm[np.ix_(idx, idx)] += o * np.outer(C[:, k], C[:, k])
    return dm


def core(H_core: np.ndarray, S: np.ndarray, nocc: int) -> np.ndarray:
    """Core-Hamiltonian guess: occupy the lowest eigenvectors of H_core C = S C e."""
    L = np.linalg.cholesky(S)
    Li = np.linalg.inv(L)
    _, v = np.linalg.eigh(Li @ H_core @ Li.T)
    C = Li.T @ v[:, :nocc]
    return 2 * C @ C.T
