"""ORACLE (test infrastructure only): RKS B3LYP nuclear gradient, FP64.

Restates pyscf_ipu/nanoDFT/nanoDFT.py:403-449 (`grad_elec`, `grad_nuc`) and
:451-477 (`grad`: the integrals it takes from PySCF) with this oracle's own
integrals in place of PySCF's:
  int2e_ip1, int1e_ipovlp, int1e_ipkin, int1e_ipnuc, int1e_iprinv
      -> cint.eri_ip1 / cint.int1e_ip (derivatives of the Rys / Obara-Saika
         integrals of cint_oracle.c through raised and lowered shells);
  NumInt().eval_ao(deriv=2) -> ao.eval_ao_deriv2;
  b3lyp -> xc.b3lyp.
Like the reference it has no grid-response term: the grid points and weights
stay where they are while the atoms move.  The gradient is therefore the
derivative of the SCF energy at a frozen grid.

Pins (tests/test_grad_oracle.py): every derivative integral against central
finite differences of the oracle's own integrals under a displaced atom; the
AO second derivatives against finite differences of the first; the total
gradient against central finite differences of the converged oracle SCF
energy on a frozen grid; the nuclear-repulsion part against its closed form.
"""
from __future__ import annotations

import numpy as np

from . import ao as oracle_ao
from . import cint
from . import xc as oracle_xc

HYB_B3LYP = 0.2
EPSILON_B3LYP = 1e-20


def aoslices(mol) -> np.ndarray:
    """(natm, 4): first shell, end shell, first AO, end AO of each atom
    (PySCF's aoslice_by_atom for a basis whose shells are listed by atom)."""
    bas = mol._bas
    width = 2 * bas[:, 1] + 1
    ao0 = np.concatenate([[0], np.cumsum(width)])
    out = np.zeros((mol.natm, 4), dtype=np.int64)
    for ia in range(mol.natm):
        sh = np.where(bas[:, 0] == ia)[0]
        if len(sh):
            out[ia] = [sh[0], sh[-1] + 1, ao0[sh[0]], ao0[sh[-1] + 1]]
        else:
            out[ia] = [0, 0, 0, 0]
    return out


def grad_nuc(charges: np.ndarray, coords: np.ndarray) -> np.ndarray:
    """d E_nuc / d R (natm, 3), nanoDFT.py:435-449."""
    natm = len(charges)
    zz = charges[:, None] * charges[None, :]
    diff = coords[None, :, :] - coords[:, None, :]          # [a, b] = R_b - R_a
    r3 = np.linalg.norm(diff, axis=2) ** 3
    r3[np.arange(natm), np.arange(natm)] = np.inf
    terms = -zz[:, :, None] * diff / r3[:, :, None]
    return terms.sum(axis=0)


def hcore_deriv(mol, h1: np.ndarray, slices: np.ndarray) -> np.ndarray:
    """(natm, 3, N, N): d H_core / d R_atom blocks as nanoDFT.py:463-470."""
    N = mol.nao_nr()
    out = np.zeros((mol.natm, 3, N, N))
    charges = mol.atom_charges()
    for ia in range(mol.natm):
        p0, p1 = slices[ia, 2:]
        vrinv = cint.int1e_ip(mol, nuc=ia)[3] * -float(charges[ia])
        vrinv[:, p0:p1] += h1[:, p0:p1]
        out[ia] = vrinv + vrinv.transpose(0, 2, 1)
    return out


def grad_elec(weights, ao, eri_ip1, s1, h1aos, slices, mask, mo_energy, mo_coeff) -> np.ndarray:
    """Electronic part (natm, 3), nanoDFT.py:403-433."""
    dm0 = 2 * (mo_coeff * mask) @ mo_coeff.T
    dme0 = 2 * (mo_coeff * mask * mo_energy) @ mo_coeff.T
    rho = np.sum((ao[:1] @ dm0) * ao, axis=2)                # (10, g)
    _, vrho, vgamma = oracle_xc.b3lyp(rho[:4], EPSILON_B3LYP)
    V = np.concatenate([vrho.reshape(1, -1) / 2, 4 * vgamma.reshape(1, -1) * rho[1:4]], axis=0)   # (4, g)
    wV = weights * V
    vmat = ao[1:4].transpose(0, 2, 1) @ np.sum(ao[:4] * wV[:, :, None], axis=0)                   # (3, N, N)
    aos = np.stack([ao[[1, 4, 5, 6]], ao[[2, 5, 7, 8]], ao[[3, 6, 8, 9]]])                        # (3, 4, g, N)
    V_xc = -vmat - np.einsum("snpi,np->spi", aos, wV).transpose(0, 2, 1) @ ao[0]
    vj = -np.einsum("sijkl,lk->sij", eri_ip1, dm0)
    vk = -np.einsum("sijkl,jk->sil", eri_ip1, dm0)
    vhf = V_xc + vj - vk * 0.5 * HYB_B3LYP
    de = np.einsum("lxij,ij->lx", h1aos, dm0)
    for k in range(len(slices)):
        p0, p1 = slices[k, 2:]
        de[k] += np.einsum("xij,ij->x", vhf[:, p0:p1], dm0[p0:p1]) * 2
        de[k] -= np.einsum("xij,ij->x", s1[:, p0:p1], dme0[p0:p1]) * 2
    return de


def grad(mol, coords, weights, mo_coeff, mo_energy) -> np.ndarray:
    """Nuclear gradient (natm, 3) in Ha/Bohr, nanoDFT.py:451-477."""
    ao = oracle_ao.eval_ao_deriv2(mol, coords)
    eri = cint.eri_ip1(mol)
    ipovlp, ipkin, ipnuc, _ = cint.int1e_ip(mol)
    s1, kin, nuc = -ipovlp, -ipkin, -ipnuc
    mask = np.ones(mol.nao_nr())
    mask[mol.nelectron // 2:] = 0
    slices = aoslices(mol)
    h1aos = hcore_deriv(mol, kin + nuc, slices)
    charges = mol.atom_charges().astype(np.float64)
    return (grad_elec(weights, ao, eri, s1, h1aos, slices, mask, mo_energy, mo_coeff)
            + grad_nuc(charges, mol.atom_coords()))
