"""B3LYP on the MI355X (drop-in for pyscf_ipu.exchange_correlation.b3lyp).

`b3lyp(rho, EPSILON_B3LYP=0)` keeps the reference signature and return triple
(pyscf_ipu/exchange_correlation/b3lyp.py:13-39): rho is (4, g) = (rho,
grad rho / 2); returns per-point (e_xc, v_rho, v_norms).  The pointwise
kernel (`pscf_b3lyp`, csrc/b3lyp.h) evaluates the functional with hand-derived
derivatives instead of autodiff.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib, device

CLIP_RHO_MIN = 1e-9
CLIP_RHO_MAX = 1e12


def b3lyp(rho, EPSILON_B3LYP=0.0):
    is_np = not isinstance(rho, torch.Tensor)
    r = rho if not is_np else torch.from_numpy(np.ascontiguousarray(rho))
    dt = torch.float64 if r.dtype == torch.float64 else torch.float32
    code = _lib.F64 if dt == torch.float64 else _lib.F32
    r = device.to_dev(r, dt).reshape(4, -1).contiguous()
    g = r.shape[1]
    exc, vrho, vn = device.empty(g, dt), device.empty(g, dt), device.empty(g, dt)
    _lib.call("pscf_b3lyp", code, device.ptr(r), g, float(EPSILON_B3LYP), device.ptr(exc), device.ptr(vrho),
              device.ptr(vn), device.stream())
    if is_np:
        return exc.cpu().numpy(), vrho.cpu().numpy(), vn.cpu().numpy()
    return exc, vrho, vn


def b3lyp_host(rho: np.ndarray, EPSILON_B3LYP=0.0):
    """The same pointwise code run on the host CPU in FP64 (no GPU)."""
    import ctypes
    r = np.ascontiguousarray(rho, dtype=np.float64).reshape(4, -1)
    g = r.shape[1]
    exc, vrho, vn = np.zeros(g), np.zeros(g), np.zeros(g)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    _lib.call("pscf_b3lyp_host", p(r), g, float(EPSILON_B3LYP), p(exc), p(vrho), p(vn))
    return exc, vrho, vn
