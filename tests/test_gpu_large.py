"""Large-molecule data paths and the precision-study GEMM on the GPU.

Every kernel is checked against a CPU restatement on the same inputs:
s8 -> s1 expansion bit-exact against the oracle's, J/K from the s1 tensor
against the dense einsum definitions, the general GEMM against numpy on the
same rounded inputs, and the whole LargeRKS SCF (6-31G, grid level 1) against
the FP64 oracle SCF."""
import numpy as np
import pytest
import torch

from oracle import cint, rks
from pyscf_ipu_amd import _lib, device, grids, molgen
from pyscf_ipu_amd.electron_repulsion import direct
from pyscf_ipu_amd.gto import build_mol
from pyscf_ipu_amd.precision import TYPES, LargeRKS

pytestmark = pytest.mark.gpu

WATER = "O 0 0 0; H 0.7586 0 0.5043; H -0.7586 0 0.5043"


def _round(x, kind):
    if kind == "f64":
        return x.astype(np.float64)
    if kind == "f32":
        return x.astype(np.float32).astype(np.float64)
    if kind == "f16":
        return x.astype(np.float16).astype(np.float64)
    u = x.astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("tin", ["bf16", "f16", "f32", "f64"])
@pytest.mark.parametrize("acc", ["f32", "f64"])
@pytest.mark.parametrize("shape", [(70, 1000, 50, 0, 0, 1), (37, 45, 3000, 0, 1, 7)])
def test_gemm_variants(hip, tin, acc, shape):
    M, N, K, ta, tb, ksplit = shape
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(K, M) if ta else (M, K))
    B = rng.normal(size=(N, K) if tb else (K, N))
    code, tdt = TYPES[tin]
    dA = torch.empty(A.shape, dtype=tdt, device="cuda")
    dB = torch.empty(B.shape, dtype=tdt, device="cuda")
    A64, B64 = device.to_dev(A), device.to_dev(B)
    st = device.stream()
    _lib.call("pscf_cast_to", code, device.ptr(A64), device.ptr(dA), A.size, st)
    _lib.call("pscf_cast_to", code, device.ptr(B64), device.ptr(dB), B.size, st)
    C = torch.zeros((M, N), dtype=torch.float64, device="cuda")
    _lib.call("pscf_gemm", code, int(acc == "f64"), ta, tb, int(ksplit > 1), M, N, K, device.ptr(dA), A.shape[1],
              device.ptr(dB), B.shape[1], device.ptr(C), N, ksplit, st)
    Ar, Br = _round(A, tin), _round(B, tin)
    ref = (Ar.T if ta else Ar) @ (Br.T if tb else Br)
    got = C.cpu().numpy()
    scale = np.sqrt(K)
    tol = 1e-12 if acc == "f64" else 2e-5
    assert np.abs(got - ref).max() < tol * scale, (tin, acc, np.abs(got - ref).max())


def test_s8_to_s1_and_jk_s1(hip):
    mol = build_mol(WATER, "6-31g")
    N = mol.nao_nr()
    s8 = direct.eri_s8(mol, "fp64")
    s1 = device.empty(N ** 4, torch.float64)
    _lib.call("pscf_eri_s8_to_s1", _lib.F64, device.ptr(s8), _lib.F64, device.ptr(s1), N, device.stream())
    ref = cint.s8_to_s1(s8.cpu().numpy(), N)
    got = s1.cpu().numpy().reshape(N, N, N, N)
    np.testing.assert_array_equal(got, ref)
    rng = np.random.default_rng(0)
    D = rng.normal(size=(N, N))
    D = D + D.T
    out = device.empty((N, N), torch.float64)
    _lib.call("pscf_jk_s1", _lib.F64, device.ptr(s1), N, device.ptr(device.to_dev(D)), 0.2, device.ptr(out),
              device.stream())
    J = np.einsum("ijkl,ji->kl", ref, D)
    K = np.einsum("ijkl,jk->il", ref, D)
    np.testing.assert_allclose(out.cpu().numpy(), J - 0.1 * K, atol=1e-11)


def _jk_s8_large(s8, code, N, D):
    ws_b = _lib.lib().pscf_jk_s8_large_workspace_bytes(N)
    ws = device.workspace(ws_b)
    out = device.empty((N, N), torch.float64)
    _lib.call("pscf_jk_s8_large", code, device.ptr(s8), N, device.ptr(device.to_dev(D)), 0.2, device.ptr(out),
              device.ptr(ws), ws_b, device.stream())
    return out.cpu().numpy()


def _sym_density(N, seed):
    D = np.random.default_rng(seed).normal(size=(N, N))
    return D + D.T


@pytest.mark.parametrize("eri", ["fp64", "fp32"])
def test_jk_s8_large_matches_oracle_631g(hip, eri):
    """J - hyb/2 K from the packed s8 ERIs with row-tiled dm and accumulator
    equals the einsum definitions over the oracle's s1 expansion of the same
    ERIs (a 5-heavy-atom molecule in 6-31G, nao about 60)."""
    mol = build_mol(molgen.plain_set(1, 5, seed=7)[0], "6-31g")
    N = mol.nao_nr()
    print(f"jk_s8_large {eri} nao {N}", flush=True)
    s8 = direct.eri_s8(mol, eri)
    code = _lib.F64 if eri == "fp64" else _lib.F32
    ref_s1 = cint.s8_to_s1(s8.double().cpu().numpy(), N)
    D = _sym_density(N, 1)
    got = _jk_s8_large(s8, code, N, D)
    J = np.einsum("ijkl,ji->kl", ref_s1, D)
    K = np.einsum("ijkl,jk->il", ref_s1, D)
    ref = J - 0.1 * K
    # FP64 accumulation of the same ERI values; the 2^-40 rounding of each
    # contribution bounds the difference
    assert np.abs(got - ref).max() < 1e-9 * max(1.0, np.abs(ref).max()), np.abs(got - ref).max()
    np.testing.assert_array_equal(got, _jk_s8_large(s8, code, N, D))


def test_jk_s8_large_equals_s1_beyond_lds_nao(hip):
    """At nao above the LDS-resident s8 kernels' limit (a 12-heavy-atom
    molecule in 6-31G) the packed form equals the s1 stream on the GPU."""
    mol = build_mol(molgen.plain_set(1, 12, seed=5)[0], "6-31g")
    N = mol.nao_nr()
    assert N > 116
    print(f"jk_s8_large vs s1 nao {N}", flush=True)
    s8 = direct.eri_s8(mol, "fp64")
    D = _sym_density(N, 2)
    got = _jk_s8_large(s8, _lib.F64, N, D)
    s1 = device.empty(N ** 4, torch.float64)
    _lib.call("pscf_eri_s8_to_s1", _lib.F64, device.ptr(s8), _lib.F64, device.ptr(s1), N, device.stream())
    out = device.empty((N, N), torch.float64)
    _lib.call("pscf_jk_s1", _lib.F64, device.ptr(s1), N, device.ptr(device.to_dev(D)), 0.2, device.ptr(out),
              device.stream())
    ref = out.cpu().numpy()
    del s1
    assert np.abs(got - ref).max() < 1e-9 * np.abs(ref).max(), np.abs(got - ref).max()


def test_large_rks_matches_oracle_631g_level1(hip):
    mol = build_mol(WATER, "6-31g")
    grid = grids.build_grid(mol, 1)
    ref = rks.scf(mol, grid=grid, its=40, conv_tol=1e-10)
    got = LargeRKS(mol, 1, "f64/f64", "fp64", grid=grids.atomic_grids(mol, 1)).run(its=40, conv_tol=1e-10)
    assert got.converged
    assert abs(got.energy - ref.energy) < 1e-7, (got.energy, ref.energy)
    assert np.abs(got.dm - ref.dm).max() < 1e-5


def test_large_rks_equals_plain_path(hip):
    """The large-molecule data paths reproduce the LDS-resident ones."""
    from pyscf_ipu_amd.nanoDFT.nanoDFT import RKS
    s = molgen.plain_set(1, 4, seed=3)[0]
    mol = build_mol(s)
    a = RKS(mol, level=0, dtype="fp64").run(its=30, conv_tol=1e-9)
    b = LargeRKS(mol, 0, "f64/f64", "fp64").run(its=30, conv_tol=1e-9)
    assert abs(a.energy - b.energy) < 1e-8
    assert abs(a.iterations - b.iterations) <= 1


def test_precision_variants_bounded(hip):
    """Lower-precision AO-grid GEMMs move the energy by bounded amounts,
    ordered by the input precision (FP32 < FP16 < BF16)."""
    mol = build_mol(WATER, "6-31g")
    grid = grids.atomic_grids(mol, 1)
    ref = LargeRKS(mol, 1, "f64/f64", "fp64", grid=grid).run(its=40, conv_tol=1e-10)
    dE = {}
    for v in ("f32/f32", "f16/f32", "bf16/f32", "bf16/f64"):
        r = LargeRKS(mol, 1, v, "fp64", grid=grid).run(its=40, conv_tol=1e-10, dm0=ref.dm)
        dE[v] = abs(r.energy - ref.energy)
    print(dE)
    assert dE["f32/f32"] < 1e-5
    assert dE["f32/f32"] < dE["f16/f32"] < dE["bf16/f32"]
    assert dE["bf16/f64"] < 1e-1
