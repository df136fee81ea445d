"""Batched form on the GPU: one launch per SCF phase over many molecules.

Each molecule's result must equal what the single-molecule path and the FP64
oracle give on the same molecule: FP64 batch within 1e-7 Ha of the oracle,
FP32 batch within 1 mHa (the metric's gate)."""
import numpy as np
import pytest

from oracle import rks
from pyscf_ipu_amd import batch, grids, molgen
from pyscf_ipu_amd.gto import METHANE, build_mol

pytestmark = pytest.mark.gpu

STRS = [METHANE] + molgen.batched_set(7, (1, 4), seed=21)


@pytest.fixture(scope="module")
def oracle_results():
    out = []
    for s in STRS:
        mol = build_mol(s)
        out.append(rks.scf(mol, grid=grids.build_grid(mol, 0), its=40, conv_tol=1e-10))
    return out


def test_batch_fp64_matches_oracle(hip, oracle_results):
    print(f"batch fp64 {len(STRS)} molecules", flush=True)
    b = batch.BatchRKS(batch.parse(STRS), dtype="fp64")
    res = b.run(max_its=40, conv_tol=1e-10, graph=False)
    for m, ref in enumerate(oracle_results):
        assert res.converged[m], m
        assert abs(res.energy[m] - ref.energy) < 1e-7, (m, res.energy[m], ref.energy)
        assert abs(res.hlgap[m] - ref.hlgap) < 1e-5
        D = b.matrix(b.D, m)
        assert np.abs(D - ref.dm).max() < 1e-4


def test_batch_fp32_within_1mha(hip, oracle_results):
    mols = batch.parse(STRS)
    eager = batch.BatchRKS(mols, dtype="fp32").run(max_its=30, conv_tol=1e-7, graph=False)
    graphed = batch.BatchRKS(mols, dtype="fp32").run(max_its=30, conv_tol=1e-7, graph=True)
    for m, ref in enumerate(oracle_results):
        assert abs(eager.energy[m] - ref.energy) < 1e-3
        assert abs(graphed.energy[m] - ref.energy) < 1e-3


@pytest.mark.parametrize("dtype", ["fp32", "fp64"])
def test_batch_runs_bit_identical(hip, dtype):
    """Two runs of the same batch give the same bits, and so do other launch
    shapes (XC and J/K workgroups per molecule, J/K workgroup size): the J/K
    atomics add contributions rounded to 2^-40, whose FP64 sums are exact in
    any order; the J/K column form's register sums and the XC partials cover
    fixed row sets and fixed 512-point chunks whatever the launch shape; every
    other reduction has a fixed order."""
    mols = batch.parse(STRS + molgen.batched_set(24, (5, 9), seed=33))
    tol = 1e-6 if dtype == "fp32" else 1e-9
    runs = []
    for xc_split, jk_split, threads in ((1, 1, 256), (1, 1, 256), (4, 4, 512), (2, 8, 256)):
        b = batch.BatchRKS(mols, dtype=dtype)
        b.setup()
        b.split, b.jk_threads = {"xc": xc_split, "jk": jk_split}, threads
        print(f"bit-identical {dtype} xc split {xc_split} jk split {jk_split} threads {threads}", flush=True)
        runs.append(b.run(max_its=30, conv_tol=tol, graph=True))
    for r in runs[1:]:
        assert np.array_equal(r.energy, runs[0].energy)
        assert np.array_equal(r.iterations, runs[0].iterations)
        assert np.array_equal(r.hlgap, runs[0].hlgap)


def test_batch_graph_replay_equals_eager_fp64(hip):
    """Replays of the captured iteration (with eager host reads in between)
    reproduce the eager sequence."""
    mols = batch.parse(STRS)
    eager = batch.BatchRKS(mols, dtype="fp64").run(max_its=30, conv_tol=1e-9, graph=False)
    graphed = batch.BatchRKS(mols, dtype="fp64").run(max_its=30, conv_tol=1e-9, graph=True, check_every=1)
    assert np.abs(graphed.energy - eager.energy).max() < 1e-9
    assert np.array_equal(graphed.iterations, eager.iterations)


@pytest.mark.parametrize("dtype,tol", [("fp64", 1e-8), ("fp32", 1e-4)])
def test_batch_restart_in_place(hip, dtype, tol):
    """Molecules restarted in place from the atomic guess (after a first run cut
    short) reach what a fresh batch reaches; the others keep their first-run
    state and stay inactive."""
    mols = batch.parse(STRS)
    fresh = batch.BatchRKS(mols, dtype=dtype).run(max_its=30, conv_tol=1e-9 if dtype == "fp64" else 1e-6)
    b = batch.BatchRKS(mols, dtype=dtype)
    first = b.run(max_its=3, conv_tol=1e-9)
    redo = np.array([0, 3, 6])
    b.restart(redo)
    again = b.run(max_its=30, conv_tol=1e-9 if dtype == "fp64" else 1e-6)
    assert again.converged[redo].all()
    assert np.abs(again.energy[redo] - fresh.energy[redo]).max() < tol
    others = np.setdiff1d(np.arange(len(mols)), redo)
    assert np.array_equal(again.iterations[others], first.iterations[others])
    # a molecule whose set-up failed (no valid L^-1) is refused, not revived
    b.ctl[5, 2] = -1.0
    with pytest.raises(ValueError):
        b.restart(np.array([2, 5]))


def test_batch_matches_plain_path(hip):
    from pyscf_ipu_amd.nanoDFT.nanoDFT import RKS
    s = STRS[3]
    b = batch.BatchRKS(batch.parse([s]), dtype="fp64").run(max_its=30, conv_tol=1e-9, graph=False)
    p = RKS(build_mol(s), level=0, dtype="fp64").run(its=30, conv_tol=1e-9)
    assert abs(b.energy[0] - p.energy) < 1e-8
    assert b.iterations[0] == p.iterations


def test_batch_split_and_tuner_fp64(hip):
    """Several workgroups per molecule (XC tiles / J/K rows split, partials
    reduced or atomically added) give the one-workgroup result, and the
    autotuner picks one of its candidates for each launch."""
    mols = batch.parse(STRS)
    ref = batch.BatchRKS(mols, dtype="fp64").run(max_its=30, conv_tol=1e-9, graph=False)
    b = batch.BatchRKS(mols, dtype="fp64")
    b.setup()
    tun = b.tune()
    assert set(tun["chosen_split"]) == {"xc", "jk"} and all(v in batch.SPLITS for v in tun["chosen_split"].values())
    assert tun["chosen_jk_threads"] in batch.JK_THREADS
    assert b.tune().get("cached")
    b.split = {"xc": 4, "jk": 8}
    b.jk_threads = 512
    got = b.run(max_its=30, conv_tol=1e-9, graph=True)
    assert np.abs(got.energy - ref.energy).max() < 1e-9
    assert np.array_equal(got.iterations, ref.iterations)


@pytest.mark.parametrize("dtype,basis,tiles,rtol", [("fp64", "sto-3g", (32,), 1e-12), ("fp32", "sto-3g", (32, 64), 2e-5),
                                                    ("fp64", "6-31g", (32,), 1e-12)])
def test_batch_xc_direct_equals_stored(hip, dtype, basis, tiles, rtol):
    """XC with the AO values evaluated inside the kernel (shell table in LDS,
    pscf_batch_xc_direct) equals XC over the AO block stored by
    pscf_batch_ao, on the atomic-guess density: molecules with 1-9 heavy atoms
    (nao classes 16-80; FP64, whose shell table fits the LDS up to nao 64, on 1-7
    heavy atoms) in STO-3G, and 6-31G molecules (six-primitive shells)."""
    import torch
    from pyscf_ipu_amd import device
    heavy = (5, 9) if dtype == "fp32" else (5, 7)
    strs = STRS + molgen.batched_set(24, heavy, seed=33) if basis == "sto-3g" else STRS[:4]
    mols = batch.parse(strs)
    got = {}
    for direct in (True, False):
        b = batch.BatchRKS(mols, basis=basis, dtype=dtype)
        b.ao_direct = direct
        b.setup()
        for tile in tiles:
            for split in (1, 2):
                print(f"xc direct={direct} {dtype} {basis} tile {tile} split {split}", flush=True)
                b._xc(device.ptr(b.ctl), split, tile)
                torch.cuda.synchronize()
                got[direct, tile, split] = (b.Vxc.double().cpu().numpy(), b.exc.cpu().numpy())
    for (direct, tile, split), (V, e) in got.items():
        V0, e0 = got[False, tiles[0], 1]
        assert np.isfinite(V).all() and np.isfinite(e).all()
        assert np.abs(V - V0).max() <= rtol * np.abs(V0).max(), (direct, tile, split)
        assert np.abs(e - e0).max() <= rtol * np.abs(e0).max(), (direct, tile, split)
    assert np.abs(got[True, tiles[0], 1][0]).max() > 0.1


def test_batch_fp32_tiles_agree(hip):
    """FP32 XC with 64-point tiles and split workgroups equals 32-point tiles."""
    mols = batch.parse(STRS)
    runs = []
    for tile, split in ((32, 1), (64, 2)):
        b = batch.BatchRKS(mols, dtype="fp32")
        b.setup()
        b.xc_tile, b.split = tile, {"xc": split, "jk": split}
        b.jk_threads = 256 * split
        runs.append(b.run(max_its=30, conv_tol=1e-6, graph=False))
    assert np.abs(runs[0].energy - runs[1].energy).max() < 1e-4  # FP32 summation-order noise at a 1e-6 stop


def _multiplicity(N):
    """2^([p==q] + [r==s] + [pq==rs]) for every element of the s8 packing,
    from numpy index arithmetic (sparse_symmetric_ERI.py:68-79)."""
    p, q = np.tril_indices(N)
    pq_diag = (p == q).astype(int)
    P = len(p)
    a, b = np.tril_indices(P)              # element (pq=a, rs=b), s8 order
    return 2.0 ** (pq_diag[a] + pq_diag[b] + (a == b))


def test_batch_nuclear_repulsion(hip):
    """pscf_batch_enuc equals the Coulomb sum over atom pairs of each molecule."""
    mols = batch.parse(STRS + molgen.batched_set(24, (5, 9), seed=33))
    b = batch.BatchRKS(mols, dtype="fp32")
    b.setup()
    got = b.d_enuc.cpu().numpy()
    for m, (Z, R) in enumerate(mols):
        ref = sum(Z[i] * Z[j] / np.linalg.norm(R[i] - R[j]) for i in range(len(Z)) for j in range(i))
        assert abs(got[m] - ref) < 1e-10 * max(1.0, ref), m


def test_batch_work_lists(hip):
    """pscf_batch_chunks: (owner, first) pairs covering range(count) of every
    owner in steps of 256, owners with zero count included, in owner order."""
    rng = np.random.default_rng(4)
    counts = rng.integers(0, 2000, size=300)
    counts[[0, 7, 299]] = 0
    got = batch.BatchRKS._work_list(counts).cpu().numpy()
    ref = [(i, s) for i, c in enumerate(counts) for s in range(0, int(c), 256)]
    assert np.array_equal(got, np.array(ref, np.int32).reshape(-1, 2))


def test_batch_eri_scale_matches_oracle(hip):
    """The batch ERI buffer holds the oracle's s8 ERIs over their multiplicity."""
    from oracle import cint
    b = batch.BatchRKS(batch.parse(STRS[:4]), dtype="fp64")
    b.setup()
    eri = b.eri.cpu().numpy()
    for m, s in enumerate(STRS[:4]):
        mol = build_mol(s)
        N = mol.nao_nr()
        ref = cint.eri_s8(mol) / _multiplicity(N)
        o = int(b.desc[m, 11])
        got = eri[o:o + ref.size]
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-13)
    # the divided-as-written form equals the raw ERIs divided in place, bit for bit
    from pyscf_ipu_amd import _lib, device
    import ctypes
    import torch
    raw = torch.empty_like(b.eri)
    works = [b._work_list(c) for c in b.eri_counts]
    arr = (ctypes.c_void_p * 6)(*[w.data_ptr() for w in works])
    cnt = (ctypes.c_int * 6)(*[w.shape[0] for w in works])
    P, st = device.ptr, device.stream()
    _lib.call("pscf_batch_eri_s8", b.code, P(b.d_desc), b.M, arr, cnt, P(b.d_atm), P(b.d_bas), P(b.d_env),
              P(b.d_shell_ao), P(b.d_shl), P(raw), st)
    _lib.call("pscf_batch_eri_scale", b.code, P(b.d_desc), b.M, b.max_nao, P(raw), st)
    torch.cuda.synchronize()
    assert torch.equal(raw, b.eri)


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
def test_batch_jk_row_and_column_forms(hip, dtype):
    """The row form (used above nao 64) and the column form give the dense
    J - (hyb/2) K of each molecule's density."""
    import torch
    from oracle import cint
    from pyscf_ipu_amd import _lib, device
    b = batch.BatchRKS(batch.parse(STRS[:4]), dtype=dtype)
    b.setup()
    P, st = device.ptr, device.stream()
    outs = []
    for cmax in (b.max_nao, 65):  # 65 sizes the launch past the column form's 64
        out = torch.zeros_like(b.dJK if dtype == "fp64" else b.Vxc)
        _lib.call("pscf_batch_jk_s8", b.code, P(b.d_desc), P(b.d_order), b.M, cmax, None, P(b.eri), P(b.Dt), 0.2,
                  P(out), 0, 1, 256, st)
        outs.append(out.double())
    torch.cuda.synchronize()
    tol = 1e-10 if dtype == "fp64" else 2e-4
    for m, s in enumerate(STRS[:4]):
        mol = build_mol(s)
        N = mol.nao_nr()
        D = b.matrix(b.D, m)
        ref = rks.dense_jk(cint.s8_to_s1(cint.eri_s8(mol), N), D)
        o = int(b.desc[m, 12])
        for out in outs:
            got = out[o:o + N * N].cpu().numpy().reshape(N, N)
            assert np.abs(got - ref).max() < tol * max(1.0, np.abs(ref).max()), (m, np.abs(got - ref).max())


def test_batch_qm1b_rows(hip, oracle_results, tmp_path):
    """QM1B rows of a batched FP64 run (energy, HOMO, LUMO, gap in eV; std of
    the last five iteration energies from the device ring) against the oracle
    SCF of the same molecules, through a Parquet round trip."""
    from pyscf_ipu_amd import qm1b
    from pyscf_ipu_amd.nanoDFT.nanoDFT import HARTREE_TO_EV
    b = batch.BatchRKS(batch.parse(STRS), dtype="fp64")
    res = b.run(max_its=40, conv_tol=1e-10, graph=True)
    rows = qm1b.rows_from_batch(b, res)
    assert len(rows) == len(STRS)
    path = str(tmp_path / "rows.parquet")
    qm1b.write_parquet(rows, path)
    df = qm1b.read_parquet(path)
    for m, ref in enumerate(oracle_results):
        nocc = build_mol(STRS[m]).nelectron // 2
        assert abs(df["energy"][m] / HARTREE_TO_EV - ref.energy) < 1e-7
        assert abs(df["homo"][m] / HARTREE_TO_EV - ref.mo_energy[nocc - 1]) < 1e-6
        assert abs(df["lumo"][m] / HARTREE_TO_EV - ref.mo_energy[nocc]) < 1e-6
        assert abs(df["y"][m] / HARTREE_TO_EV - ref.hlgap) < 1e-6
        if ref.iterations >= 5 and res.iterations[m] == ref.iterations:
            assert abs(df["std"][m] - np.std(ref.energies[-5:, 0] * HARTREE_TO_EV)) < 1e-5
        assert df["N"][m] == build_mol(STRS[m]).nao_nr()
