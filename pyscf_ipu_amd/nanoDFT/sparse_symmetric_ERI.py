"""Sparse 8-fold-symmetric J/K on the MI355X (drop-in for pyscf_ipu.nanoDFT.sparse_symmetric_ERI).

Kept names (pyscf_ipu/nanoDFT/sparse_symmetric_ERI.py):
  get_i_j(val)                                    :13-16
  num_repetitions_fast(ij, kl)                    :68-79
  indices_func(i, j, k, l, symmetry, N)           :81-84 (== compute_indices.cpp:54-95)
  sparse_symmetric_einsum(nonzero_distinct_ERI, nonzero_indices, dm, backend)  :86-137
The index helpers are exact integer arithmetic on the host; the contraction
runs in `pscf_jk_sparse` (all 16 gather/scale/scatter patterns in one pass).
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib, device

HYB_B3LYP = 0.2


def _isqrt_floor(x: np.ndarray) -> np.ndarray:
    """floor(sqrt(x)) for non-negative int64 arrays, exact."""
    x = np.asarray(x, dtype=np.int64)
    r = np.floor(np.sqrt(x.astype(np.float64))).astype(np.int64)
    r -= (r * r > x).astype(np.int64)
    r += ((r + 1) * (r + 1) <= x).astype(np.int64)
    return r


def get_i_j(val):
    """Decode a packed lower-triangle index c = i(i+1)/2 + j into (i, j)."""
    val = np.asarray(val).astype(np.int64)
    i = (_isqrt_floor(1 + 8 * val) - 1) // 2
    j = val - i * (i + 1) // 2
    return i.astype(np.uint64), j.astype(np.uint64)


def num_repetitions_fast(ij, kl):
    """Multiplicity 2^([i==j] + [k==l] + [ij==kl]) of each unique ERI."""
    i, j = get_i_j(ij)
    k, l = get_i_j(kl)
    return 2 ** (np.equal(i, j).astype(np.uint64) + np.equal(k, l).astype(np.uint64)
                 + (1 - ((1 - np.equal(k, i) * np.equal(l, j)) * (1 - np.equal(k, j) * np.equal(l, i))).astype(np.uint64)))


def indices_func(i, j, k, l, symmetry, N):
    """Flat index of pattern `symmetry` (0-31): J gather 0-7, J scatter 8-15,
    K gather 16-23, K scatter 24-31."""
    return np.array([i * N + j, j * N + i, i * N + j, j * N + i, k * N + l, l * N + k, k * N + l, l * N + k,
                     k * N + l, k * N + l, l * N + k, l * N + k, i * N + j, i * N + j, j * N + i, j * N + i,
                     k * N + j, k * N + i, l * N + j, l * N + i, i * N + l, i * N + k, j * N + l, j * N + k,
                     i * N + l, j * N + l, i * N + k, j * N + k, k * N + j, l * N + j, k * N + i, l * N + i])[symmetry]


def prepare_sparse(eri_s8: np.ndarray, threshold: float = 0.0):
    """Nonzero s8 ERIs divided by multiplicity and their int16 (i, j, k, l), as
    nanoDFT.py:285-313 prepares them (without the device/batch reshape)."""
    e = np.array(eri_s8, copy=True)
    e[np.abs(e) <= threshold] = 0.0
    nz = np.nonzero(e)[0].astype(np.uint64)
    v = e[nz]
    ij, kl = get_i_j(nz)
    rep = num_repetitions_fast(ij, kl)
    i, j = get_i_j(ij)
    k, l = get_i_j(kl)
    idx = np.stack([i, j, k, l], axis=1).astype(np.int16)
    return v / rep, idx


def _as_int16_indices(nonzero_indices):
    if isinstance(nonzero_indices, torch.Tensor):
        t = nonzero_indices
        if t.dtype == torch.float16:
            t = t.view(torch.int16)
        return t.reshape(-1, 4).to(torch.int16)
    a = np.asarray(nonzero_indices)
    if a.dtype == np.float16:
        a = a.view(np.int16)
    return a.reshape(-1, 4).astype(np.int16)


def sparse_symmetric_einsum(nonzero_distinct_ERI, nonzero_indices, dm, backend="mi355x", hyb=HYB_B3LYP):
    """dJK = J - hyb/2 K (flat, N*N) from multiplicity-divided unique ERIs.

    Accepts the reference's shapes ((batches, n) values, (batches, n, 4) int16
    indices possibly bit-cast to float16, dm flat or (N, N)); the arithmetic type
    follows `nonzero_distinct_ERI` (float32 or float64).  Returns a device
    tensor, or a numpy array when `dm` is a numpy array."""
    is_np = not isinstance(dm, torch.Tensor)
    vals = nonzero_distinct_ERI if isinstance(nonzero_distinct_ERI, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(nonzero_distinct_ERI))
    dt = torch.float64 if vals.dtype == torch.float64 else torch.float32
    code = _lib.F64 if dt == torch.float64 else _lib.F32
    vals = device.to_dev(vals.reshape(-1), dt)
    idx = device.to_dev(_as_int16_indices(nonzero_indices))
    dmt = device.to_dev(dm if isinstance(dm, torch.Tensor) else np.asarray(dm), dt).reshape(-1)
    N = int(round(np.sqrt(dmt.numel())))
    if N * N != dmt.numel():
        raise ValueError("dm must have N*N elements")
    nnz = vals.numel()
    if idx.shape[0] != nnz:
        raise ValueError("ERI values and indices differ in length")
    out = device.empty(N * N, dt)
    wsb = _lib.lib().pscf_jk_sparse_workspace_bytes(code, N, nnz)
    ws = device.workspace(wsb)
    _lib.call("pscf_jk_sparse", code, device.ptr(vals), device.ptr(idx), nnz, device.ptr(dmt), N, float(hyb),
              device.ptr(out), device.ptr(ws), wsb, device.stream())
    return out.cpu().numpy() if is_np else out


def jk_s8(eri_s8: torch.Tensor, dm: torch.Tensor, hyb: float = HYB_B3LYP, out=None, ws=None) -> torch.Tensor:
    """dJK (N, N) from the packed s8 array (raw ERIs), symmetric dm; indices
    regenerated on the device."""
    dt = eri_s8.dtype
    code = _lib.F64 if dt == torch.float64 else _lib.F32
    N = dm.shape[0]
    if out is None:
        out = device.empty((N, N), dt)
    wsb = _lib.lib().pscf_jk_s8_workspace_bytes(code, N)
    if ws is None or ws.numel() < wsb:
        ws = device.workspace(wsb)
    _lib.call("pscf_jk_s8", code, device.ptr(eri_s8), N, device.ptr(dm), float(hyb), device.ptr(out),
              device.ptr(ws), wsb, device.stream())
    return out
