"""Integer index lists (bit-exact) and grid construction, on the CPU."""
import math

import numpy as np
import pytest

from oracle import indices as oidx
from pyscf_ipu_amd import grids, gto, molgen
from pyscf_ipu_amd.electron_repulsion import direct
from pyscf_ipu_amd.nanoDFT import sparse_symmetric_ERI as ssE


@pytest.mark.parametrize("N", [1, 5, 9, 37, 70])
def test_get_i_j_matches_bruteforce(N):
    P = N * (N + 1) // 2
    vals = np.arange(P)
    i, j = ssE.get_i_j(vals)
    bi, bj = oidx.get_i_j_bruteforce(vals, N)
    assert np.array_equal(i.astype(np.int64), bi) and np.array_equal(j.astype(np.int64), bj)


def test_get_i_j_large_indices_exact():
    # s8 compound indices for nao = 200 (config with 6-31G): c up to ~2e8
    N = 200
    P = N * (N + 1) // 2
    c = np.array([0, 1, 2, P * (P + 1) // 2 - 1, 123456789, 98765432])
    i, j = ssE.get_i_j(c)
    i, j = i.astype(np.int64), j.astype(np.int64)
    assert np.all(j <= i) and np.all(i * (i + 1) // 2 + j == c)


def test_num_repetitions_matches_bruteforce():
    N = 7
    P = N * (N + 1) // 2
    ij, kl = np.meshgrid(np.arange(P), np.arange(P), indexing="ij")
    m = ij >= kl
    ij, kl = ij[m], kl[m]
    rep = ssE.num_repetitions_fast(ij, kl)
    i, j = ssE.get_i_j(ij)
    k, l = ssE.get_i_j(kl)
    for a in range(len(ij)):
        assert rep[a] == oidx.multiplicity_bruteforce(int(i[a]), int(j[a]), int(k[a]), int(l[a]))


def test_index_table_all_32_patterns():
    N = 11
    rng = np.random.default_rng(0)
    for _ in range(200):
        i, j, k, l = rng.integers(0, N, 4)
        ref = oidx.index_table(i, j, k, l, N)
        for sym in range(32):
            assert ssE.indices_func(i, j, k, l, sym, N) == ref[sym]


@pytest.mark.parametrize("nbas", [1, 2, 7, 13])
def test_shell_quartet_list_matches_reference_loop(nbas):
    got = direct.quartet_list(nbas)
    ref = np.array(oidx.shell_quartets_bruteforce(nbas), dtype=np.int32).reshape(-1, 4)
    assert np.array_equal(got, ref)


def test_prepare_electron_repulsion_integrals_structure():
    mol = gto.build_mol(gto.METHANE)  # 7 shells
    out = direct.prepare_electron_repulsion_integrals(mol)
    input_floats, input_ints, tuple_ijkl, shapes, sizes, counts, indxs, indxs_inv, num_calls = out
    assert input_floats.dtype == np.float32 and input_floats.shape == (1, len(mol._env))
    assert input_ints.shape == (1, 6 + 8 + 6 * mol.natm + 8 * mol.nbas)
    assert list(input_ints[0, :6]) == [81, 81, mol.natm, mol.nbas, len(mol._env), 8]
    assert num_calls == len(oidx.shell_quartets_bruteforce(mol.nbas)) == sum(counts)
    assert sorted(sizes) == list(sizes) and set(sizes) <= {1, 3, 9, 27, 81}
    assert np.array_equal(indxs[indxs_inv], np.arange(num_calls))
    flat = [q for grp in tuple_ijkl for q in grp]
    ref = oidx.shell_quartets_bruteforce(mol.nbas)
    assert [tuple(ref[k]) for k in indxs] == flat
    # duplicated quartet classes of the reference filter (i*j >= k*l): count them
    canon = {tuple(sorted([tuple(sorted(q[:2])), tuple(sorted(q[2:]))])) for q in ref}
    assert len(ref) - len(canon) == 24  # 7 shells


def test_shell_pairs_by_class_cover_all_pairs():
    mol = gto.build_mol(molgen.plain_set(1, 4, seed=3)[0])
    pairs, counts = direct.shell_pairs_by_class(mol)
    assert sum(counts) == mol.nbas * (mol.nbas + 1) // 2
    l = mol._bas[:, 1]
    assert {tuple(sorted(p)) for p in pairs.tolist()} == {(j, i) for i in range(mol.nbas) for j in range(i + 1)}
    s0, s1 = counts[0], counts[0] + counts[1]
    assert np.all(l[pairs[:s0]].sum(1) == 0) and np.all(l[pairs[s0:s1, 0]] == 1) and np.all(l[pairs[s0:s1, 1]] == 0)
    assert np.all(l[pairs[s1:]].sum(1) == 2)


def _sph_harm_real(l, m, v):
    """Unnormalised real solid harmonics via explicit low-order polynomials."""
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    return {(1, 0): z, (2, 0): 3 * z * z - 1, (2, 2): x * x - y * y, (3, 1): x * (5 * z * z - 1),
            (4, 0): 35 * z ** 4 - 30 * z * z + 3, (4, 4): x ** 4 - 6 * x * x * y * y + y ** 4,
            (5, 3): x * (x * x - 3 * y * y) * (9 * z * z - 1), (6, 2): (x * x - y * y) * (33 * z ** 4 - 18 * z * z + 1),
            (7, 1): x * (429 * z ** 6 - 495 * z ** 4 + 135 * z * z - 5)}[(l, m)]


@pytest.mark.parametrize("npts,order", [(50, 11), (74, 13), (86, 15), (110, 17)])
def test_lebedev_exactness(npts, order):
    g = grids.lebedev(npts)
    assert g.shape == (npts, 4)
    assert np.allclose(np.linalg.norm(g[:, :3], axis=1), 1.0, atol=1e-14)
    assert abs(g[:, 3].sum() - 1.0) < 1e-14
    # every monomial x^a y^b z^c of degree <= order integrates exactly
    for a in range(order + 1):
        for b in range(order + 1 - a):
            for c in range(order + 1 - a - b):
                val = np.sum(g[:, 3] * g[:, 0] ** a * g[:, 1] ** b * g[:, 2] ** c)
                if a % 2 or b % 2 or c % 2:
                    exact = 0.0
                else:
                    df = lambda n: math.prod(range(n, 0, -2)) if n > 0 else 1  # noqa: E731
                    exact = df(a - 1) * df(b - 1) * df(c - 1) / df(a + b + c + 1)
                assert abs(val - exact) < 1e-13, (a, b, c)
    for (l, m) in [(1, 0), (2, 0), (2, 2), (3, 1), (4, 0), (4, 4), (5, 3), (6, 2), (7, 1)]:
        if l <= order:
            assert abs(np.sum(g[:, 3] * _sph_harm_real(l, m, g[:, :3]))) < 1e-12


def test_treutler_radial_quadrature():
    # int_0^inf r^2 exp(-a r^2) dr = sqrt(pi)/(4 a^1.5)
    r, dr = grids.treutler_ahlrichs(75)
    for a in [0.3, 1.0, 5.0]:
        assert abs(np.sum(r ** 2 * dr * np.exp(-a * r * r)) - math.sqrt(math.pi) / (4 * a ** 1.5)) < 1e-8


def test_becke_weights_partition_of_unity():
    mol = gto.build_mol("O 0 0 0; C 1.3 0 0; H -0.6 0.8 0; H 1.9 0.9 0.1")
    g = grids.atomic_grids(mol, 0)
    # sum over owners of the partition functions at one point is 1: evaluate each
    # atom's weight function at a common set of points
    pts = np.random.default_rng(0).normal(size=(50, 3))
    tot = np.zeros(len(pts))
    for ia in range(mol.natm):
        tot += grids.becke_partition(pts, mol.atom_coords(), mol.atom_charges(), np.full(len(pts), ia))
    assert np.allclose(tot, 1.0, atol=1e-13)


def test_grid_integrates_electron_count():
    from oracle import ao, cint, xc
    from pyscf_ipu_amd import guess
    mol = gto.build_mol(gto.METHANE)
    g = grids.build_grid(mol, 1)
    S, _, _ = cint.int1e(mol)
    dm = guess.minao(mol, S)
    a = ao.eval_ao_deriv1(mol, g.coords)
    _, _, rho = xc.exchange_correlation(dm, a, g.weights)
    # Becke-partitioned level-1 grid: electron count to ~1e-3 (level 0: ~5e-3)
    assert abs(np.sum(rho[0] * g.weights) - np.sum(dm * S)) < 2e-3
    assert abs(np.sum(dm * S) - mol.nelectron) < 1e-10


def test_grid_sizes_level0():
    # H: 10 radial shells, C/N/O: 15, angular orders pruned by NWChem's scheme
    for z, nrad in [(1, 10), (6, 15), (7, 15), (8, 15)]:
        c, v = grids.atomic_grid(z, 0)
        r = np.unique(np.round(np.linalg.norm(c, axis=1), 12))
        assert len(r) == nrad


@pytest.mark.parametrize("mol_str", [gto.METHANE, "O 0 0 0; H 0 0.9 0.3; H 0 -0.9 0.3",
                                     "C 0 0 0; N 1.4 0 0; O 0 1.3 0; H -1 0 0; H 0 0 1"])
def test_einsum_do_lists_match_reference_loop(mol_str):
    """prepare_einsum_inputs (vectorised first-occurrence flags) equals the
    reference's dictionary loop (direct.py:1029-1110), bit for bit, and the
    applied index images cover every ordered shell quartet exactly once."""
    mol = gto.build_mol(mol_str)
    ind, do, N, n = direct.prepare_einsum_inputs(mol)
    ref_ind, ref_do = oidx.einsum_inputs_bruteforce(mol._bas)
    assert np.array_equal(np.array(ind), ref_ind) and np.array_equal(np.array(do), ref_do)
    assert N == mol.nao_nr() and n == len(ref_ind)
    assert int(np.sum(do)) == mol.nbas ** 4
