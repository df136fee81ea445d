"""Parity at the benchmark's own sizes, against committed FP64 oracle results.

* Plain and batched forms on 8 of the 32 seeded 9-heavy-atom STO-3G molecules
  of the plain benchmark (nao 54-65; tests/golden/plain_sto3g_l0.json): the
  XC tile pipelines for padded nao 64 and 80, the J/K row form above nao 64 and
  the eigensolver at these sizes.  FP64 within 1e-7 Ha and the HOMO-LUMO gap
  within 1e-5 Ha of the oracle; FP32 within the 1 mHa gate.
* The large-molecule path (6-31G, grid level 1, nao 148-152;
  tests/golden/large_631g_l1.json) in FP64 and with FP32 AO-grid GEMMs.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PLAIN = json.load(open(os.path.join(HERE, "golden", "plain_sto3g_l0.json")))["molecules"]
PICK = [0, 3, 7, 11, 15, 19, 23, 27]
LARGE_PATH = os.path.join(HERE, "golden", "large_631g_l1.json")
# The FP32 forms stop at |dE| < 1e-6 Ha (plain: 1e-5), above their rounding
# noise.  The HOMO-LUMO gap converges more slowly than the energy: molecule 3 of
# the set stops with a gap 1.6 mHa from the converged one, and the FP64 oracle
# stopped at the same |dE| < 1e-6 gives that same gap (0.04188 against 0.04348
# Ha fully converged).  The gap error is stated, the energy is the gate.
FP32_GAP_TOL = 3e-3


def _gold(key):
    return np.array([PLAIN[i][key] for i in PICK])


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
def test_plain_at_benchmark_size(hip, dtype):
    from pyscf_ipu_amd.gto import build_mol
    from pyscf_ipu_amd.nanoDFT.nanoDFT import run_molecules
    mols = [build_mol(PLAIN[i]["mol_str"]) for i in PICK]
    print(f"plain {dtype} nao {[m.nao_nr() for m in mols]}", flush=True)
    # the plain FP32 form builds J/K from the whole FP32 density each iteration
    # (the reference's --float32 mode), so its energy carries ~1e-6 Ha of
    # rounding noise from step to step: it stops at 1e-5
    res = run_molecules(mols, streams=4, dtype=dtype, its=60, conv_tol=1e-9 if dtype == "fp64" else 1e-5)
    E = np.array([r.energy for r in res])
    gap = np.array([r.hlgaps[-1] for r in res])
    assert all(r.converged for r in res)
    if dtype == "fp64":
        assert np.abs(E - _gold("energy")).max() < 1e-7, np.abs(E - _gold("energy"))
        assert np.abs(gap - _gold("hlgap")).max() < 1e-5
    else:
        assert np.abs(E - _gold("energy")).max() < 1e-3, np.abs(E - _gold("energy"))
        assert np.abs(gap - _gold("hlgap")).max() < FP32_GAP_TOL


@pytest.mark.parametrize("dtype", ["fp64", "fp32"])
def test_batch_at_benchmark_size(hip, dtype):
    from pyscf_ipu_amd import batch
    strs = [PLAIN[i]["mol_str"] for i in PICK]
    print(f"batch {dtype} {len(strs)} nine-heavy-atom molecules", flush=True)
    b = batch.BatchRKS(batch.parse(strs), dtype=dtype)
    r = b.run(max_its=40, conv_tol=1e-9 if dtype == "fp64" else 1e-6, graph=True)
    assert r.converged.all()
    dE = np.abs(r.energy - _gold("energy"))
    dg = np.abs(r.hlgap - _gold("hlgap"))
    if dtype == "fp64":
        assert dE.max() < 1e-7, dE
        assert dg.max() < 1e-5
    else:
        assert dE.max() < 1e-3, dE
        assert dg.max() < FP32_GAP_TOL


@pytest.mark.skipif(not os.path.exists(LARGE_PATH), reason="large-molecule golden file not generated")
@pytest.mark.parametrize("gemm,eri,tol", [("f64/f64", "fp64", 1e-7), ("f32/f32", "fp64", 1e-3),
                                          ("f32/f32", "fp32", 1e-3)])
def test_large_631g_at_study_size(hip, gemm, eri, tol):
    import torch
    from pyscf_ipu_amd import grids
    from pyscf_ipu_amd.gto import build_mol
    from pyscf_ipu_amd.precision import LargeRKS
    gold = json.load(open(LARGE_PATH))
    for m in gold["molecules"]:
        mol = build_mol(m["mol_str"], gold["basis"])
        print(f"large {gemm} eri {eri} nao {mol.nao_nr()}", flush=True)
        grid = grids.atomic_grids(mol, gold["grid_level"])
        # FP32 GEMMs: stop at |dE| < 1e-6, above their energy noise (the stop
        # at 1e-7 sat at that noise and missed the 60 iterations now and then)
        r = LargeRKS(mol, gold["grid_level"], gemm, eri, grid=grid).run(its=60, conv_tol=1e-9 if tol < 1e-6
                                                                          else 1e-6)
        assert r.converged
        assert abs(r.energy - m["energy"]) < tol, (r.energy, m["energy"])
        assert abs(r.hlgaps[-1] - m["hlgap"]) < (1e-5 if tol < 1e-6 else 1e-4)
        del r
        torch.cuda.empty_cache()
