"""ORACLE (test infrastructure only): AO values and gradients on grid points.

Restates what the reference takes from PySCF's `mol.eval_gto('GTOval_sph_deriv1',
coords, 4)` (pyscf_ipu/nanoDFT/nanoDFT.py:219-220): a (4, g, N) array of
[phi, dphi/dx, dphi/dy, dphi/dz] for spherical s/p AOs (p in x, y, z order).
Pinned by finite differences of phi and by int phi_i phi_j w = S_ij on the
molecular grid (tests/test_oracle_grid_xc.py).
"""
from __future__ import annotations

import numpy as np


def eval_ao_deriv1(mol, coords: np.ndarray) -> np.ndarray:
    coords = np.asarray(coords, dtype=np.float64)
    g = len(coords)
    out = np.zeros((4, g, mol.nao_nr()))
    col = 0
    for (R, l, exps, coefs) in mol.shell_arrays():
        d = coords - R
        r2 = np.einsum("gi,gi->g", d, d)
        e = np.exp(-np.outer(r2, exps))          # (g, nprim)
        rad = e @ coefs                          # sum_k c_k exp(-a_k r^2)
        drad = -2.0 * (e * exps) @ coefs         # d/d(r^2) * 2 -> multiply by d_x
        if l == 0:
            out[0, :, col] = rad
            for x in range(3):
                out[1 + x, :, col] = drad * d[:, x]
            col += 1
        else:
            for a in range(3):
                out[0, :, col + a] = d[:, a] * rad
                for x in range(3):
                    out[1 + x, :, col + a] = (rad if a == x else 0.0) + d[:, a] * drad * d[:, x]
            col += 3
    return out
