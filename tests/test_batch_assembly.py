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
        assert b.ngrid[m] == grids.atomic_grids(mol, 0).size
    # offsets are contiguous
    nao = b.nao
    assert np.array_equal(b.desc[:, 12], np.cumsum(nao * nao) - nao * nao)
    P = nao * (nao + 1) // 2
    assert b.sizes["eri"] == int((P * (P + 1) // 2).sum())


def test_eri_counts_cover_all_quartets():
    """The six classes' quartet counts of each molecule add up to all its
    unique shell quartets (the work lists cover range(count) per molecule)."""
    strs = molgen.batched_set(25, (1, 6), seed=5)
    b = batch.BatchRKS(batch.parse(strs), dtype="fp32")
    ns, npp = b.desc[:, 6], b.desc[:, 7]
    tot = np.zeros(b.M, np.int64)
    for (X, Y), nq in zip(batch.CLASSES, b.eri_counts):
        nx, ny = batch.class_pairs(X, ns, npp), batch.class_pairs(Y, ns, npp)
        assert np.array_equal(nq, nx * (nx + 1) // 2 if X == Y else nx * ny)
        tot += nq
    nsh = b.desc[:, 1]
    npair = nsh * (nsh + 1) // 2
    assert np.array_equal(tot, npair * (npair + 1) // 2)


def test_qm1b_rows_roundtrip_parquet(tmp_path):
    """QM1B-schema rows (qm1b/README.md:7-18) from plain-form results: unit
    conversions, column types and a Parquet round trip."""
    from pyscf_ipu_amd import qm1b
    from pyscf_ipu_amd.gto import METHANE, build_mol
    from pyscf_ipu_amd.nanoDFT.nanoDFT import HARTREE_TO_EV, SCFResult
    mol = build_mol(METHANE)
    N = mol.nao_nr()
    E = np.array([-40.1, -40.2, -40.25, -40.26, -40.261, -40.2611])
    energies = np.zeros((6, 5))
    energies[:, 0] = E
    mo = np.linspace(-10, 1, N)
    r = SCFResult(E[-1], energies, np.zeros(6), mo, np.eye(N), np.eye(N), 6, True)
    rows = qm1b.rows_from_plain([mol], [r], smiles=["C"])
    assert len(rows) == 1
    row = rows[0]
    assert row["atoms"] == "CHHHH" and row["z"] == [6, 1, 1, 1, 1] and row["N"] == N
    assert abs(row["energy"] - E[-1] * HARTREE_TO_EV) < 1e-9
    assert abs(row["homo"] - mo[4] * HARTREE_TO_EV) < 1e-9 and abs(row["lumo"] - mo[5] * HARTREE_TO_EV) < 1e-9
    assert abs(row["y"] - (mo[5] - mo[4]) * HARTREE_TO_EV) < 1e-9
    assert abs(row["std"] - np.std(E[-5:] * HARTREE_TO_EV)) < 1e-12
    assert np.allclose(np.array(row["pos"]).reshape(-1, 3), mol.atom_coords())
    assert qm1b.formula([6, 1, 1, 1, 1, 8]) == "CH4O"
    p = tmp_path / "qm1b.parquet"
    qm1b.write_parquet(rows * 3, str(p))
    df = qm1b.read_parquet(str(p))
    assert list(df.columns) == list(qm1b.COLUMNS) and len(df) == 3
    assert df["smile"][0] == "C" and list(df["z"][2]) == [6, 1, 1, 1, 1]
    assert abs(df["energy"][1] - row["energy"]) == 0.0
    # a molecule that did not converge is left out unless asked for
    r.converged = False
    assert qm1b.rows_from_plain([mol], [r]) == []
    assert qm1b.rows_from_plain([mol], [r], only_converged=False)[0]["smile"] == "CH4"
