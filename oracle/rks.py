"""ORACLE (test infrastructure only): FP64 restatement of nanoDFT's RKS SCF.

Follows pyscf_ipu/nanoDFT/nanoDFT.py:
  energy            :20-38   E = sum D*H_core + 1/2 sum D*dJK + E_xc + E_nuc
  nanoDFT_iteration :40-102  H = H_core + dJK + V_xc -> DIIS -> eigh(L^-1 H L^-T) -> D = 2 C_occ C_occ^T
  exchange_corr.    :105-123 (oracle/xc.py)
  get_JK            :125-139 dJK = J - hyb/2 K (sparse 16-pass form, oracle/cint.c)
  init tensors      :208-257 S, T, V, L^-1 = inv(chol(S)), occupancy, DIIS(9) buffers
  DIIS              :338-364 errvec = S D H - (S D H)^T, ring buffer, pinv(B)[0]
  hlgap             :366-368
with one addition the reference lacks: a stop when |E_i - E_(i-1)| < conv_tol
(the reference always runs `its` iterations, nanoDFT.py:157).

`xc="hf"` switches to Hartree-Fock (dJK = J - K/2, no V_xc), used only to pin
integrals + SCF against textbook STO-3G Hartree-Fock energies.

Parity pins: HF total energies of Szabo & Ostlund Table 3.13 (STO-3G); the SCF
fixed-point conditions (commutator FDS - SDF -> 0, idempotency DSD = 2D,
tr(DS) = N_electrons).  B3LYP total energies against PySCF are unpinned here.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import ao as ao_mod
from . import cint, xc

HYB_B3LYP = 0.2
DIIS_ITERS = 9


def s8_sparse(mol, eri_s8: Optional[np.ndarray] = None):
    """Nonzero s8 ERIs divided by their multiplicity, with (i,j,k,l) int16
    indices, as nanoDFT.py:285-313 prepares them (FP64 values kept)."""
    if eri_s8 is None:
        eri_s8 = cint.eri_s8(mol)
    nz = np.nonzero(eri_s8)[0].astype(np.int64)
    v = eri_s8[nz]
    ij = ((np.sqrt(1 + 8 * nz.astype(np.float64)) - 1) // 2).astype(np.int64)
    kl = nz - ij * (ij + 1) // 2
    i = ((np.sqrt(1 + 8 * ij.astype(np.float64)) - 1) // 2).astype(np.int64)
    j = ij - i * (i + 1) // 2
    k = ((np.sqrt(1 + 8 * kl.astype(np.float64)) - 1) // 2).astype(np.int64)
    l = kl - k * (k + 1) // 2
    rep = 2.0 ** ((i == j).astype(int) + (k == l).astype(int) + (ij == kl).astype(int))
    idx = np.stack([i, j, k, l], axis=1).astype(np.int16)
    return v / rep, idx


def dense_jk(eri_s1: np.ndarray, dm: np.ndarray, hyb: float = HYB_B3LYP) -> np.ndarray:
    """Dense definition (nanoDFT.py:130-132)."""
    J = np.einsum("ijkl,ji->kl", eri_s1, dm)
    K = np.einsum("ijkl,jk->il", eri_s1, dm)
    return J - K / 2 * hyb


def energy(dm, H_core, dJK, E_xc, E_nuc):
    E_core = float(np.sum(dm * H_core))
    E_JK = float(np.sum(dm * dJK))
    return np.array([E_core + E_JK / 2 + E_xc + E_nuc, E_core, E_JK / 2, E_xc, E_nuc])


def pinv_sym(a: np.ndarray, rcond: Optional[float] = None) -> np.ndarray:
    """Pseudo-inverse of a symmetric matrix (== SVD pinv) with numpy/JAX's
    default relative cutoff 10 * max(shape) * eps."""
    if rcond is None:
        rcond = 10 * max(a.shape) * np.finfo(np.float64).eps
    w, v = np.linalg.eigh(a)
    cut = rcond * np.max(np.abs(w))
    winv = np.where(np.abs(w) > cut, 1.0 / np.where(w == 0, 1, w), 0.0)
    return (v * winv) @ v.T


@dataclass
class DIISState:
    V: np.ndarray
    H: np.ndarray
    B: np.ndarray


def diis_init(N: int) -> DIISState:
    B = np.zeros((DIIS_ITERS + 1, DIIS_ITERS + 1))
    B[0, 1:] = B[1:, 0] = 1
    return DIISState(np.zeros((DIIS_ITERS, N * N)), np.zeros((DIIS_ITERS, N * N)), B)


def diis(i: int, H: np.ndarray, dm: np.ndarray, S: np.ndarray, st: DIISState):
    head = i % DIIS_ITERS
    sdf = S @ dm @ H
    err = (sdf - sdf.T).reshape(-1)
    st.V[head] = err
    st.H[head] = H.reshape(-1)
    mask = (np.arange(DIIS_ITERS) < min(i + 1, DIIS_ITERS)).astype(np.float64)
    tmps = (st.V @ err) * mask
    st.B[head + 1, 1:] = tmps
    st.B[1:, head + 1] = tmps
    m_ = np.concatenate([[1.0], mask])
    c = pinv_sym(st.B * m_[:, None] * m_[None, :])[0]
    return (c[1:] @ st.H).reshape(H.shape)


@dataclass
class SCFResult:
    energy: float
    energies: np.ndarray          # per iteration [E, E_core, E_JK, E_xc, E_nuc]
    dm: np.ndarray
    mo_energy: np.ndarray
    mo_coeff: np.ndarray
    hlgap: float
    iterations: int
    converged: bool
    H: np.ndarray = field(repr=False, default=None)


def scf(mol, grid=None, its: int = 20, conv_tol: float = 1e-8, xc_name: str = "b3lyp", dm0=None,
        eri_s8=None, level: int = 0, diis_on: bool = True) -> SCFResult:
    from pyscf_ipu_amd import grids as grids_mod
    from pyscf_ipu_amd import guess as guess_mod

    N = mol.nao_nr()
    nocc = mol.nelectron // 2
    S, T, V = cint.int1e(mol)
    H_core = T + V
    L_inv = np.linalg.inv(np.linalg.cholesky(S))
    E_nuc = mol.energy_nuc()
    occ = np.concatenate([np.ones(nocc), np.zeros(N - nocc)])
    hyb = HYB_B3LYP if xc_name == "b3lyp" else 1.0
    eri_v, eri_idx = s8_sparse(mol, eri_s8)
    if xc_name == "b3lyp":
        if grid is None:
            grid = grids_mod.build_grid(mol, level)
        aos = ao_mod.eval_ao_deriv1(mol, grid.coords)

    def fock_parts(dm):
        if xc_name == "b3lyp":
            E_xc, V_xc, _ = xc.exchange_correlation(dm, aos, grid.weights)
        else:
            E_xc, V_xc = 0.0, np.zeros_like(dm)
        return E_xc, V_xc, cint.sparse_jk(eri_v, eri_idx, dm, hyb)

    dm = guess_mod.minao(mol, S) if dm0 is None else dm0
    E_xc, V_xc, dJK = fock_parts(dm)
    st = diis_init(N)
    E_prev = energy(dm, H_core, dJK, E_xc, E_nuc)[0]
    log = []
    converged = False
    H = None
    for i in range(its):
        H = H_core + dJK + V_xc
        if diis_on:
            H = diis(i, H, dm, S, st)
        _, v = np.linalg.eigh(L_inv @ H @ L_inv.T)
        C = L_inv.T @ v
        dm = (C * occ * 2) @ C.T
        E_xc, V_xc, dJK = fock_parts(dm)
        e = energy(dm, H_core, dJK, E_xc, E_nuc)
        log.append(e)
        if abs(e[0] - E_prev) < conv_tol:
            converged = True
            break
        E_prev = e[0]
    mo_energy, v = np.linalg.eigh(L_inv @ H @ L_inv.T)
    mo_coeff = L_inv.T @ v
    gap = abs(mo_energy[nocc] - mo_energy[nocc - 1])
    log = np.array(log)
    return SCFResult(float(log[-1, 0]), log, dm, mo_energy, mo_coeff, float(gap), len(log), converged, H)
