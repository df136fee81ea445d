"""Device-buffer plumbing over PyTorch-ROCm (allocation, streams, pointers).

PyTorch only provides HIP memory and the current stream here; all arithmetic
of the hot path runs in the library's own kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

DTYPES = {"fp64": (torch.float64, _lib.F64, np.float64), "fp32": (torch.float32, _lib.F32, np.float32)}


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.HIPError("no HIP device visible: the MI355X path has no CPU fallback")
    _lib.lib()
    return torch.device("cuda", torch.cuda.current_device())


def ptr(t) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(t.data_ptr())


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_dev(a, dtype=None) -> torch.Tensor:
    dev = require_gpu()
    if isinstance(a, torch.Tensor):
        t = a.to(dev)
    else:
        t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    if dtype is not None:
        t = t.to(dtype)
    return t.contiguous()


def upload_many(*arrays):
    """Host arrays to the device in one copy: packed (16-byte aligned) into one
    host buffer, copied once, returned as typed views of the device buffer."""
    arrs = [np.ascontiguousarray(a) for a in arrays]
    offs, total = [], 0
    for a in arrs:
        total = (total + 15) & ~15
        offs.append(total)
        total += a.nbytes
    buf = np.zeros(max(total, 16), np.uint8)
    for a, o in zip(arrs, offs):
        buf[o:o + a.nbytes] = a.reshape(-1).view(np.uint8)
    d = torch.from_numpy(buf).to(require_gpu())
    return [d[o:o + a.nbytes].view(torch.from_numpy(a[:0].reshape(-1)).dtype).reshape(a.shape)
            for a, o in zip(arrs, offs)]


def empty(shape, dtype) -> torch.Tensor:
    return torch.empty(shape, dtype=dtype, device=require_gpu())


def zeros(shape, dtype) -> torch.Tensor:
    return torch.zeros(shape, dtype=dtype, device=require_gpu())


def workspace(nbytes: int) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 8), dtype=torch.uint8, device=require_gpu())
