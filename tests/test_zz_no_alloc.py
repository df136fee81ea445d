"""pscf_jk_s8_noalloc, the allocation-free variant of pscf_jk_s8, through the
C-ABI: two calls in a row (a 9-heavy-atom molecule, then methane, smaller),
each against pscf_jk_s8 with a caller workspace on the same inputs and against
the oracle's J - hyb/2 K; plus the size limit, which the host rejects before
any launch."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import cint, rks
from pyscf_ipu_amd import _lib, device, molgen
from pyscf_ipu_amd.gto import METHANE, build_mol

pytestmark = pytest.mark.gpu

BIG = molgen.plain_set(1, 9, seed=3)[0]


def _case(mol, dtype, seed):
    N = mol.nao_nr()
    rng = np.random.default_rng(seed)
    dm = rng.normal(size=(N, N))
    dm = dm + dm.T
    e8 = cint.eri_s8(mol)
    ref = cint.sparse_jk(*rks.s8_sparse(mol, e8), dm)
    tdt = device.DTYPES[dtype][0]
    return N, device.to_dev(e8, tdt), device.to_dev(dm, tdt), ref, tdt


@pytest.mark.parametrize("dtype,tol", [("fp64", 1e-11), ("fp32", 1e-5)])
def test_jk_s8_noalloc_two_calls(hip, dtype, tol):
    code = _lib.F64 if dtype == "fp64" else _lib.F32
    for mol_str, seed in ((BIG, 1), (METHANE, 2)):
        mol = build_mol(mol_str)
        N, eri, dm, ref, tdt = _case(mol, dtype, seed)
        print(f"noalloc {dtype} nao={N}", flush=True)
        out_na = device.empty((N, N), tdt)
        _lib.call("pscf_jk_s8_noalloc", code, device.ptr(eri), N, device.ptr(dm), 0.2, device.ptr(out_na),
                  device.stream())
        wsb = _lib.lib().pscf_jk_s8_workspace_bytes(code, N)
        ws = device.workspace(wsb)
        out_ws = device.empty((N, N), tdt)
        _lib.call("pscf_jk_s8", code, device.ptr(eri), N, device.ptr(dm), 0.2, device.ptr(out_ws), device.ptr(ws),
                  wsb, device.stream())
        torch.cuda.synchronize()
        a, b = out_na.double().cpu().numpy(), out_ws.double().cpu().numpy()
        scale = max(1.0, np.abs(ref).max())
        # same kernels and partial layout; the LDS accumulation order may differ
        assert np.abs(a - b).max() < tol * scale
        assert np.abs(a - ref).max() < tol * scale
        assert np.array_equal(a, a.T)


def test_jk_s8_noalloc_rejects_oversize(hip):
    L = _lib.lib()
    cap = L.pscf_jk_s8_noalloc_max_bytes()
    N = 9
    while L.pscf_jk_s8_workspace_bytes(_lib.F64, N) <= cap:
        N += 8
    # the size check comes before any pointer is used or kernel launched
    rc = L.pscf_jk_s8_noalloc(_lib.F64, ctypes.c_void_p(16), N, ctypes.c_void_p(16), 0.2, ctypes.c_void_p(16), None)
    assert rc == -1
    assert b"scratch" in L.pscf_last_error()
