"""CPU-side checks of the C-ABI boundary: the library loads and exports every
symbol include/pyscf_mi355x.h declares; host-evaluable entry points work."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pyscf_mi355x.h")).read()
    return sorted(set(re.findall(r"\b(pscf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from pyscf_ipu_amd import _lib
    L = _lib.lib()
    for name in _declared():
        assert hasattr(L, name), name
    assert set(_declared()) == set(_lib.symbols())
    assert L.pscf_abi_version() == 1


def test_invalid_arguments_return_error_without_gpu():
    from pyscf_ipu_amd import _lib
    L = _lib.lib()
    rc = L.pscf_jk_s8(7, None, 4, None, 0.2, None, None, 0, None)
    assert rc == -1
    assert b"bad dtype" in L.pscf_last_error()
    assert L.pscf_b3lyp(0, None, -1, 0.0, None, None, None, None) == -1


def test_host_b3lyp_matches_oracle_autograd():
    """The kernel's hand-derived derivatives against torch autograd of the
    reference's formulas, on the host (same source code as the device)."""
    from oracle import xc
    from pyscf_ipu_amd.exchange_correlation.b3lyp import b3lyp_host
    rng = np.random.default_rng(0)
    g = 3000
    rho = np.zeros((4, g))
    rho[0] = 10 ** rng.uniform(-12, 3, g)
    rho[1:4] = rng.normal(size=(3, g)) * rho[0] * 10 ** rng.uniform(-3, 1, g)
    rho[0, :4] = [-1e-3, 0.0, 2e13, 1.0]  # clipping edges
    a = xc.b3lyp(rho, 1e-20)
    b = b3lyp_host(rho, 1e-20)
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y) / (np.abs(x) + 1e-300)) < 1e-11


def test_pool_refuses_after_gpu_use(monkeypatch):
    """The pool's workers are started by fork + exec, which must not follow GPU
    use in the starting process."""
    import torch
    from pyscf_ipu_amd.nanoDFT.pool import MoleculePool
    monkeypatch.setattr(torch.cuda, "is_initialized", lambda: True)
    with pytest.raises(RuntimeError, match="before this process uses the GPU"):
        MoleculePool(workers=1)


def test_pool_refuses_under_a_profiler(monkeypatch):
    """A profiler's preloaded library initialises the GPU before the program
    starts, so no workers may be forked from such a process."""
    from pyscf_ipu_amd.nanoDFT import pool
    monkeypatch.setenv("ROCPROF_KERNEL_TRACE", "1")
    assert "profiler" in pool.start_blocker()
    with pytest.raises(RuntimeError, match="profiler"):
        pool.MoleculePool(workers=1)
    monkeypatch.delenv("ROCPROF_KERNEL_TRACE")
    monkeypatch.setenv("LD_PRELOAD", "/opt/rocm/lib/librocprofiler-sdk-tool.so")
    assert "profiler" in pool.start_blocker()


def test_pool_reports_a_worker_that_cannot_start():
    """Without a GPU a worker fails to start: the pool raises with the worker's
    error, leaves no process behind and refuses later use."""
    import torch
    from pyscf_ipu_amd.nanoDFT import pool
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present: workers would start")
    with pytest.raises(RuntimeError, match="failed to start"):
        pool.MoleculePool(workers=1, start_timeout=120)
    p = object.__new__(pool.MoleculePool)
    p._procs = []
    with pytest.raises(RuntimeError, match="closed"):
        p.run([])


def test_bench_reads_committed_traffic_profile():
    """bench.py's roofline `traffic` comes from the committed counter summary:
    the dominant plain kernel's row parses to HBM bytes per dispatch, and the
    summary states the copy-kernel calibration it was scaled by."""
    import importlib.util
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    b = bench._profile_traffic("scf_step_kernel<double>")
    assert isinstance(b, int) and 1e5 < b < 1e8
    assert bench._profile_traffic("no_such_kernel") is None
    text = open(bench.TRAFFIC_PROFILE).read()
    assert "copy-calibration" in text and "FETCH_SIZE is scaled by" in text
