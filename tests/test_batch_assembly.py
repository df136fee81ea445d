"""Host assembly of the batched form against per-molecule Mole objects (CPU)."""
import numpy as np

from pyscf_ipu_amd import batch, grids, gto, molgen


def test_batch_descriptors_match_molecules():
    strs = molgen.batched_set(40, (1, 5), seed=11) + [gto.METHANE]
    mols = batch.parse(strs)
    b = batch.BatchRKS(mols, dtype="fp32")
    for m, s in enumerate(strs):
        mol = gto.build_mol(s)
        d = b.desc[m]
        assert d[0] == mol.natm and d[1] == mol.nbas and d[2] == mol.nao_nr() and d[3] == mol.nelectron // 2
        bas = b.bas[d[5]:d[5] + d[1]]
        # same shells (l, exponents, coefficients, centre) in the same order
        for ib in range(mol.nbas):
            lb = mol._bas[ib]
            assert bas[ib, 1] == lb[1] and bas[ib, 2] == lb[2]
            assert np.array_equal(b.env[bas[ib, 5]:bas[ib, 5] + lb[2]], mol._env[lb[5]:lb[5] + lb[2]])
            assert np.array_equal(b.env[bas[ib, 6]:bas[ib, 6] + lb[2]], mol._env[lb[6]:lb[6] + lb[2]])
            c = b.atm[bas[ib, 0], 1]
            assert np.allclose(b.env[c:c + 3], mol.atom_coords()[lb[0]])
            assert b.shell_ao[d[5] + ib] == mol.ao_loc_nr()[ib]
        # class-sorted shell list: s shells then p shells of this molecule
        sl = b.shl_sorted[d[8]:d[8] + d[1]]
        assert sorted(sl) == list(range(d[5], d[5] + d[1]))
        ls = b.bas[sl, 1]
        assert np.all(ls[:d[6]] == 0) and np.all(ls[d[6]:] == 1)
        assert abs(b.e_nuc[m] - mol.energy_nuc()) < 1e-10
        assert b.ngrid[m] == grids.atomic_grids(mol, 0).size
    # offsets are contiguous
    nao = b.nao
    assert np.array_equal(b.desc[:, 12], np.cumsum(nao * nao) - nao * nao)
    P = nao * (nao + 1) // 2
    assert b.sizes["eri"] == int((P * (P + 1) // 2).sum())


def test_eri_work_covers_all_quartets():
    strs = molgen.batched_set(25, (1, 6), seed=5)
    b = batch.BatchRKS(batch.parse(strs), dtype="fp32")
    ns, npp = b.desc[:, 6], b.desc[:, 7]
    tot = 0
    for (X, Y), w in zip(batch.CLASSES, b.eri_work):
        nx, ny = batch.class_pairs(X, ns, npp), batch.class_pairs(Y, ns, npp)
        nq = nx * (nx + 1) // 2 if X == Y else nx * ny
        # every quartet index of every molecule is covered by exactly one chunk
        for m in range(b.M):
            starts = np.sort(w[w[:, 0] == m, 1])
            assert np.array_equal(starts, np.arange(0, nq[m], 256))
        tot += nq.sum()
    nsh = b.desc[:, 1]
    npair = nsh * (nsh + 1) // 2
    assert tot == int((npair * (npair + 1) // 2).sum())
