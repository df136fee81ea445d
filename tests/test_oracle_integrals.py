"""Pins the FP64 oracle's integrals and SCF before it is trusted as the checker.

Known answers come from the reference's own tests (pyscf_ipu test/test_integrals.py)
and from textbook tables; a second ERI algorithm (THO/MD) cross-checks the Rys path.
"""
import math

import mpmath
import numpy as np
import pytest

from oracle import cint, md_eri, rks
from pyscf_ipu_amd import gto


def _h2():
    return gto.Mole([("H", (0, 0, 0)), ("H", (1.4, 0, 0))], unit="Bohr")


def test_boys_closed_forms():
    for m in range(7):
        assert cint.boys(6, 0.0)[m] == pytest.approx(1.0 / (2 * m + 1), rel=1e-15)
    for T in [1e-3, 0.5, 3.0, 11.9, 12.1, 30.0, 80.0]:
        f = cint.boys(5, T)
        assert f[0] == pytest.approx(0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T)), rel=1e-14)
        for m in range(6):
            ref = float(mpmath.quad(lambda t: t ** (2 * m) * mpmath.e ** (-T * t * t), [0, 1]))
            assert f[m] == pytest.approx(ref, rel=1e-13)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
@pytest.mark.parametrize("T", [0.0, 1e-6, 0.3, 2.0, 9.0, 25.0, 60.0])
def test_rys_rule_reproduces_moments(n, T):
    u, w = cint.rys(n, T)
    assert np.all((u > 0) & (u < 1)) and np.all(w > 0)
    for k in range(2 * n):
        ref = float(mpmath.quad(lambda t: t ** (2 * k) * mpmath.e ** (-T * t * t), [0, 1]))
        assert np.sum(w * u ** k) == pytest.approx(ref, rel=1e-12, abs=1e-300)


def test_h2_sto3g_one_electron_szabo_ostlund():
    # test/test_integrals.py:24-32 (overlap 0.6648 is for the single 1.24-scaled primitive;
    # the contracted STO-3G value of Szabo-Ostlund eq. 3.229 is 0.6593), :48-60, :81-99
    S, T, V = cint.int1e(_h2())
    assert S[0, 1] == pytest.approx(0.6593, abs=1e-4)
    np.testing.assert_allclose(T, [[0.7600, 0.2365], [0.2365, 0.7600]], atol=1e-4)
    np.testing.assert_allclose(V, np.array([[-1.2266, -0.5974], [-0.5974, -0.6538]])
                               + np.array([[-0.6538, -0.5974], [-0.5974, -1.2266]]), atol=1e-4)


def test_h2_sto3g_eri_szabo_ostlund():
    # test/test_integrals.py:119-133 (Szabo-Ostlund eq. 3.235)
    e = cint.s8_to_s1(cint.eri_s8(_h2()), 2)
    assert e[0, 0, 0, 0] == pytest.approx(0.7746, abs=1e-4)
    assert e[1, 1, 1, 1] == pytest.approx(0.7746, abs=1e-4)
    assert e[0, 0, 1, 1] == pytest.approx(0.5697, abs=1e-4)
    assert e[1, 0, 0, 0] == pytest.approx(0.4441, abs=1e-4)
    assert e[0, 1, 1, 1] == pytest.approx(0.4441, abs=1e-4)
    assert e[1, 0, 1, 0] == pytest.approx(0.2970, abs=1e-4)


def test_pyquante_primitives():
    # test/test_integrals.py:51 (T = 1.5), :85 (V = -1.595769), :114-117 (ERIs)
    # PyQuante's primitives are normalised: N_s^2 = (2a/pi)^(3/2), N_px^2 = 4a N_s^2 at a = 1.
    ns2 = (2 / math.pi) ** 1.5
    np2 = 4 * ns2
    assert ns2 * ns2 * md_eri.eri_primitive(1.0, (0, 0, 0), (0, 0, 0), 1.0, (0, 0, 0), (0, 0, 0), 1.0, (0, 0, 0),
                                            (0, 0, 0), 1.0, (0, 0, 0), (0, 0, 0)) == pytest.approx(1.128379, abs=1e-6)
    assert ns2 * np2 * md_eri.eri_primitive(1.0, (0, 0, 0), (0, 0, 0), 1.0, (0, 0, 0), (0, 0, 0), 1.0, (0, 0, 0),
                                            (1, 0, 0), 1.0, (0, 0, 0), (1, 0, 0)) == pytest.approx(0.940316, abs=1e-6)
    # Unnormalised unit-exponent s primitive at the origin through the Rys oracle.
    m = gto.Mole([("H", (0, 0, 0))], unit="Bohr")
    m._env = m._env.copy()
    pe, pc = m._bas[0, 5], m._bas[0, 6]
    m._bas = m._bas.copy()
    m._bas[0, 2] = 1
    m._env[pe] = 1.0
    m._env[pc] = 1.0 / gto.ANGULAR_FACTOR[0]
    S, T, V = cint.int1e(m)
    assert T[0, 0] / S[0, 0] == pytest.approx(1.5, abs=1e-12)
    assert V[0, 0] * ns2 == pytest.approx(-1.595769, abs=1e-6)


@pytest.mark.parametrize("atoms", [
    [("O", (0, 0, 0.1165)), ("H", (0, 0.7694, -0.4661)), ("H", (0, -0.7694, -0.4661))],
    [("N", (0, 0, 0)), ("C", (1.2, 0.3, -0.2)), ("H", (-0.5, 0.9, 0.1))],
])
def test_rys_matches_independent_tho_md(atoms):
    m = gto.Mole(atoms)
    rys = cint.s8_to_s1(cint.eri_s8(m), m.nao_nr())
    md = md_eri.eri_dense(m)
    assert np.abs(rys - md).max() < 1e-12


def _bohr_mol(atoms):
    return gto.Mole(atoms, unit="Bohr")


def test_hf_sto3g_textbook_energies():
    # Szabo & Ostlund, Table 3.13: STO-3G SCF total energies at experimental geometries.
    r = 2.050 / math.sqrt(3)
    ch4 = _bohr_mol([("C", (0, 0, 0))] + [("H", (a * r, b * r, c * r))
                                          for a, b, c in [(1, 1, 1), (1, -1, -1), (-1, 1, -1), (-1, -1, 1)]])
    th, R = math.radians(104.52) / 2, 1.809
    h2o = _bohr_mol([("O", (0, 0, 0)), ("H", (R * math.sin(th), 0, R * math.cos(th))),
                     ("H", (-R * math.sin(th), 0, R * math.cos(th)))])
    cases = [(_h2(), -1.117), (ch4, -39.727), (h2o, -74.963),
             (_bohr_mol([("N", (0, 0, 0)), ("N", (2.074, 0, 0))]), -107.496),
             (_bohr_mol([("C", (0, 0, 0)), ("O", (2.132, 0, 0))]), -111.225)]
    for mol, ref in cases:
        res = rks.scf(mol, xc_name="hf", its=60, conv_tol=1e-10)
        assert res.converged
        assert res.energy == pytest.approx(ref, abs=1.5e-3)


def test_sparse_jk_matches_dense_definition():
    m = gto.Mole("N 0 0 0; C 1.3 0.2 0; O -1.1 0.5 0.3; H 0 1 -0.8")
    N = m.nao_nr()
    e8 = cint.eri_s8(m)
    rng = np.random.default_rng(0)
    dm = rng.normal(size=(N, N))
    dm = dm + dm.T
    v, idx = rks.s8_sparse(m, e8)
    sparse = cint.sparse_jk(v, idx, dm)
    dense = rks.dense_jk(cint.s8_to_s1(e8, N), dm)
    assert np.abs(sparse - dense).max() < 1e-12
