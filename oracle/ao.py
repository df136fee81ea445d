"""ORACLE (test infrastructure only): AO values and gradients on grid points.

Restates what the reference takes from PySCF's `mol.eval_gto('GTOval_sph_deriv1',
coords, 4)` (pyscf_ipu/nanoDFT/nanoDFT.py:219-220): a (4, g, N) array of
[phi, dphi/dx, dphi/dy, dphi/dz] for spherical s/p AOs (p in x, y, z order),
and, for the nuclear gradient, NumInt().eval_ao(mol, coords, deriv=2)
(nanoDFT.py:425): (10, g, N) with the second derivatives xx, xy, xz, yy, yz,
zz appended.  Pinned by finite differences of phi and of the first
derivatives (tests/test_indices_grids.py, tests/test_grad_oracle.py) and by
int phi_i phi_j w = S_ij on the molecular grid.
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


# second-derivative components in PySCF's order after the first derivatives
_D2 = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def eval_ao_deriv2(mol, coords: np.ndarray) -> np.ndarray:
    """(10, g, N): value, gradient, then the six second derivatives.
    For a primitive sum f(r^2) = sum_k c_k exp(-a_k r^2) with d = r - R:
    d_x f = f1 d_x, d_x d_y f = f2 d_x d_y + f1 delta_xy, where
    f1 = -2 sum c a e, f2 = 4 sum c a^2 e."""
    coords = np.asarray(coords, dtype=np.float64)
    g = len(coords)
    out = np.zeros((10, g, mol.nao_nr()))
    col = 0
    for (R, l, exps, coefs) in mol.shell_arrays():
        d = coords - R
        r2 = np.einsum("gi,gi->g", d, d)
        e = np.exp(-np.outer(r2, exps))
        f0 = e @ coefs
        f1 = -2.0 * (e * exps) @ coefs
        f2 = 4.0 * (e * exps ** 2) @ coefs
        if l == 0:
            out[0, :, col] = f0
            for x in range(3):
                out[1 + x, :, col] = f1 * d[:, x]
            for k, (x, y) in enumerate(_D2):
                out[4 + k, :, col] = f2 * d[:, x] * d[:, y] + (f1 if x == y else 0.0)
            col += 1
        else:
            for a in range(3):  # phi_a = d_a f
                out[0, :, col + a] = d[:, a] * f0
                for x in range(3):
                    out[1 + x, :, col + a] = (f0 if a == x else 0.0) + d[:, a] * f1 * d[:, x]
                for k, (x, y) in enumerate(_D2):
                    # d_x d_y (d_a f) = delta_ax d_y f + delta_ay d_x f + d_a d_x d_y f
                    v = d[:, a] * (f2 * d[:, x] * d[:, y] + (f1 if x == y else 0.0))
                    if a == x:
                        v = v + f1 * d[:, y]
                    if a == y:
                        v = v + f1 * d[:, x]
                    out[4 + k, :, col + a] = v
            col += 3
    return out
