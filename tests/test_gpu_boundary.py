"""The kept reference signatures on the GPU, through the C-ABI, against the
FP64 oracle: ipu_einsum (block-wise direct J/K), get_JK(dense_ERI=True), the
nanoDFT(mol, opts) entry point (sparse and dense ERI), and the set-up kernels
of the plain path (overlap Cholesky / L^-1 and the atomic-density guess)."""
import numpy as np
import pytest
import torch

from oracle import cint, rks
from pyscf_ipu_amd import device, grids, guess, molgen
from pyscf_ipu_amd.gto import METHANE, build_mol

pytestmark = pytest.mark.gpu

MOLS = {"methane": METHANE, "gen3": molgen.plain_set(1, 3, seed=7)[0]}


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", list(MOLS))
def test_ipu_einsum_against_dense_definition(hip, name):
    from pyscf_ipu_amd.electron_repulsion import direct
    mol = build_mol(MOLS[name])
    N = mol.nao_nr()
    print(f"ipu_einsum {name} nao={N}", flush=True)
    s1 = cint.s8_to_s1(cint.eri_s8(mol), N)
    rng = np.random.default_rng(2)
    dm = rng.normal(size=(N, N))  # not symmetric: every index image must be the right one
    vj_ref = np.einsum("ijkl,ji->kl", s1, dm)
    vk_ref = np.einsum("ijkl,jk->il", s1, dm)
    out3 = direct.electron_repulsion_integrals(None, None, mol)
    vj, vk = direct.ipu_einsum(out3, dm.astype(np.float32), mol)
    scale = max(np.abs(vj_ref).max(), np.abs(vk_ref).max())
    # FP32 blocks and FP32 LDS accumulation: ~1e-6 relative to the largest entry
    assert np.abs(vj - vj_ref).max() < 2e-5 * scale
    assert np.abs(vk - vk_ref).max() < 2e-5 * scale
    # FP64 blocks (the same quartets, FP64 kernel) to near machine precision
    q = np.concatenate([np.asarray(t, dtype=np.int32).reshape(-1, 4)
                        for t in direct.prepare_electron_repulsion_integrals(mol)[2]])
    flat = direct.eri_quartets(mol, q, "fp64")
    parts, s = [], 0
    for o in out3:
        parts.append(flat[s:s + o.numel()].view(o.shape))
        s += o.numel()
    vj64, vk64 = direct.ipu_einsum(parts, dm, mol)
    assert np.abs(vj64 - vj_ref).max() < 1e-11 * scale
    assert np.abs(vk64 - vk_ref).max() < 1e-11 * scale


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_get_JK_dense(hip, dtype, tol):
    from pyscf_ipu_amd.nanoDFT.nanoDFT import get_JK
    mol = build_mol(MOLS["gen3"])
    N = mol.nao_nr()
    s1 = cint.s8_to_s1(cint.eri_s8(mol), N)
    rng = np.random.default_rng(4)
    dm = rng.normal(size=(N, N))
    ref = rks.dense_jk(s1, dm)
    got = get_JK(dm.astype(dtype), s1[None].astype(dtype), dense_ERI=True, backend="mi355x")
    assert got.shape == (N, N) and got.dtype == dtype
    assert np.abs(got - ref).max() < tol * max(1.0, np.abs(ref).max())
    # device tensors in, device tensor out
    got_t = get_JK(torch.from_numpy(dm).cuda(), torch.from_numpy(s1).cuda(), True, "mi355x")
    assert isinstance(got_t, torch.Tensor) and got_t.is_cuda
    assert np.abs(_np(got_t) - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("dense", [False, True])
def test_nanoDFT_entry_point_methane(hip, dense):
    """nanoDFT(mol, nanoDFT_options(...)) as the reference calls it: `its`
    iterations (no convergence stop), energies (its, 5) in eV."""
    from pyscf_ipu_amd.nanoDFT.nanoDFT import HARTREE_TO_EV, nanoDFT, nanoDFT_options
    opts, mol_str = nanoDFT_options(its=12, mol_str="methane", dense_ERI=dense)
    mol = build_mol(mol_str, opts.basis)
    E, (logged, gaps, mo_e, mo_c, coords, weights) = nanoDFT(mol, opts)
    assert E.shape == (12, 5) and gaps.shape == (12,)
    grid = grids.Grid(coords, weights, None, None)
    ref = rks.scf(mol, grid=grid, its=12, conv_tol=-1.0)
    assert ref.energies.shape[0] == 12
    # FP64, same iterations from the same guess: per-iteration energies agree
    assert np.abs(E / HARTREE_TO_EV - ref.energies).max() < 1e-7
    assert abs(gaps[-1] / HARTREE_TO_EV - ref.hlgap) < 1e-7
    # mo_coeff are S-orthonormal eigenvectors of the final Fock matrix
    S = cint.int1e(mol)[0]
    assert np.abs(mo_c.T @ S @ mo_c - np.eye(mol.nao_nr())).max() < 1e-9


def test_plain_setup_cholesky_and_guess(hip):
    """The plain path's set-up kernels on one molecule: L^-1 = inv(chol(S))
    (nanoDFT.py:239) and the atomic-density guess (guess.minao on the host:
    same occupations and per-atom Gram-Schmidt), with its invariants
    D = D^T and tr(D S) = N_electrons."""
    from pyscf_ipu_amd.nanoDFT.nanoDFT import RKS
    for s in (METHANE, MOLS["gen3"]):
        mol = build_mol(s)
        r = RKS(mol, level=0, dtype="fp64")
        r.setup()
        S = _np(r.S)
        Li = _np(r.L_inv)
        ref = np.linalg.inv(np.linalg.cholesky(S))
        assert np.abs(Li - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())
        assert np.abs(Li @ S @ Li.T - np.eye(len(S))).max() < 1e-10
        D = _np(r.D)
        Dh = guess.minao(mol, S)
        assert np.abs(D - Dh).max() < 1e-11
        assert np.abs(D - D.T).max() == 0.0 or np.abs(D - D.T).max() < 1e-14
        assert abs(np.sum(D * S) - mol.nelectron) < 1e-10
