"""Pins of the gradient oracle (oracle/grad.py) on the CPU.

* Derivative integrals (cint.eri_ip1, cint.int1e_ip) against central finite
  differences of the oracle's own ERIs and one-electron integrals when one
  atom is displaced: a dropped raise/lower term, a wrong sign or a missing
  operator derivative fails these.
* AO second derivatives against finite differences of the first derivatives.
* The whole gradient against central finite differences of the converged
  oracle SCF energy with the grid held fixed (the reference's gradient has no
  grid-response term, nanoDFT.py:451-477), and the nuclear part against its
  closed form.
"""
import numpy as np
import pytest

from oracle import ao as oracle_ao
from oracle import cint, rks
from oracle import grad as oracle_grad
from pyscf_ipu_amd import grids
from pyscf_ipu_amd.gto import BOHR, build_mol

WATER = "O 0.0 0.05 0.1173; H 0.02 0.7572 -0.4692; H 0.0 -0.7572 -0.4592"
HCN = "H 0 0 -1.066; C 0 0.03 0; N 0.02 0 1.156"


def _displaced(mol_str, atom, axis, h_bohr):
    parts = [p.split() for p in mol_str.split(";")]
    parts[atom][1 + axis] = repr(float(parts[atom][1 + axis]) + h_bohr * BOHR)
    return "; ".join(" ".join(p) for p in parts)


def _fd(fn, mol_str, atom, axis, h=1e-4):
    return (fn(build_mol(_displaced(mol_str, atom, axis, h))) - fn(build_mol(_displaced(mol_str, atom, axis, -h)))) / (2 * h)


def _on_atom(mol, atom):
    sl = oracle_grad.aoslices(mol)[atom]
    m = np.zeros(mol.nao_nr(), bool)
    m[sl[2]:sl[3]] = True
    return m


@pytest.mark.parametrize("atom", [0, 1])
def test_eri_ip1_against_finite_differences(atom):
    mol = build_mol(WATER)
    N = mol.nao_nr()
    ip1 = cint.eri_ip1(mol)
    on = _on_atom(mol, atom)
    for axis in range(3):
        M = ip1[axis] * on[:, None, None, None]
        # d(ij|kl)/dR_A: each index on atom A contributes minus its ip1 term
        ana = -(M + M.transpose(1, 0, 2, 3) + M.transpose(2, 3, 0, 1) + M.transpose(2, 3, 1, 0))
        fd = _fd(lambda m: cint.s8_to_s1(cint.eri_s8(m), N), WATER, atom, axis)
        assert np.abs(ana - fd).max() < 2e-7 * max(1.0, np.abs(ana).max()), (axis, np.abs(ana - fd).max())


@pytest.mark.parametrize("mol_str", [WATER, HCN])
def test_int1e_ip_against_finite_differences(mol_str):
    mol = build_mol(mol_str)
    Z = mol.atom_charges()
    for atom in range(mol.natm):
        ipovlp, ipkin, ipnuc, iprinv = cint.int1e_ip(mol, nuc=atom)
        on = _on_atom(mol, atom)
        for axis in range(3):
            fdS, fdT, fdV = (_fd(lambda m, k=k: cint.int1e(m)[k], mol_str, atom, axis) for k in range(3))
            sym = lambda A: A + A.T
            dS = -sym(ipovlp[axis] * on[:, None])
            dT = -sym(ipkin[axis] * on[:, None])
            # basis-function derivative plus the derivative of -Z_A / |r - R_A| itself
            dV = -sym(ipnuc[axis] * on[:, None]) - Z[atom] * sym(iprinv[axis])
            for ana, fd in ((dS, fdS), (dT, fdT), (dV, fdV)):
                assert np.abs(ana - fd).max() < 2e-7 * max(1.0, np.abs(ana).max())


def test_ao_deriv2_against_finite_differences():
    mol = build_mol(HCN)
    rng = np.random.default_rng(3)
    pts = rng.normal(scale=1.2, size=(40, 3))
    d2 = oracle_ao.eval_ao_deriv2(mol, pts)
    d1 = oracle_ao.eval_ao_deriv1(mol, pts)
    assert np.abs(d2[:4] - d1).max() < 1e-14
    h = 1e-5
    comps = {(0, 0): 4, (0, 1): 5, (0, 2): 6, (1, 1): 7, (1, 2): 8, (2, 2): 9}
    for (x, y), k in comps.items():
        e = np.zeros(3)
        e[y] = h
        fd = (oracle_ao.eval_ao_deriv1(mol, pts + e)[1 + x] - oracle_ao.eval_ao_deriv1(mol, pts - e)[1 + x]) / (2 * h)
        assert np.abs(d2[k] - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())


def test_grad_nuc_closed_form():
    mol = build_mol(WATER)
    R, Z = mol.atom_coords(), mol.atom_charges().astype(float)
    g = oracle_grad.grad_nuc(Z, R)
    h = 1e-5
    for a in range(mol.natm):
        for x in range(3):
            def e_nuc(dx):
                Rd = R.copy()
                Rd[a, x] += dx
                return sum(Z[i] * Z[j] / np.linalg.norm(Rd[i] - Rd[j]) for i in range(len(Z)) for j in range(i))
            assert abs(g[a, x] - (e_nuc(h) - e_nuc(-h)) / (2 * h)) < 1e-8


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g"])
def test_gradient_against_energy_finite_differences(basis):
    """dE/dR of the converged B3LYP energy on a frozen grid, for every atom
    and axis of a distorted water molecule."""
    mol = build_mol(WATER, basis)
    grid = grids.build_grid(mol, 0)
    ref = rks.scf(mol, grid=grid, its=60, conv_tol=1e-12)
    assert ref.converged
    g = oracle_grad.grad(mol, grid.coords, grid.weights, ref.mo_coeff, ref.mo_energy)
    dm0 = ref.dm
    h = 1e-4
    for atom in range(mol.natm):
        for axis in range(3):
            e = []
            for s in (h, -h):
                m = build_mol(_displaced(WATER, atom, axis, s), basis)
                e.append(rks.scf(m, grid=grid, its=60, conv_tol=1e-12, dm0=dm0).energy)
            fd = (e[0] - e[1]) / (2 * h)
            # measured agreement 1.6e-8 Ha/Bohr (FD truncation and SCF convergence)
            assert abs(g[atom, axis] - fd) < 1e-7, (atom, axis, g[atom, axis], fd)
