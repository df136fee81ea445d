"""ctypes binding of libpyscf_mi355x.so (the C-ABI declared in include/pyscf_mi355x.h).

There is no fallback: if the library is missing or no HIP device is present,
the calls that need it raise.  Build with `python -m pyscf_ipu_amd.build`.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_size_t, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpyscf_mi355x.so")

F64, F32 = 0, 1

_SIGS = {
    "pscf_last_error": (ctypes.c_char_p, []),
    "pscf_abi_version": (c_int, []),
    "pscf_eri_s8": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, POINTER(c_int), c_int,
                            c_void_p, c_void_p]),
    "pscf_eri_quartets": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int,
                                  c_void_p, c_void_p]),
    "pscf_int1e": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                           c_void_p]),
    "pscf_jk_sparse_workspace_bytes": (c_size_t, [c_int, c_int, c_int64]),
    "pscf_jk_sparse": (c_int, [c_int, c_void_p, c_void_p, c_int64, c_void_p, c_int, c_double, c_void_p, c_void_p,
                               c_size_t, c_void_p]),
    "pscf_jk_s8_workspace_bytes": (c_size_t, [c_int, c_int]),
    "pscf_jk_blocks_workspace_bytes": (c_size_t, [c_int, c_int, c_int64]),
    "pscf_jk_blocks": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                               c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pscf_jk_s8": (c_int, [c_int, c_void_p, c_int, c_void_p, c_double, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pscf_jk_s8_noalloc_max_bytes": (c_size_t, []),
    "pscf_jk_s8_noalloc": (c_int, [c_int, c_void_p, c_int, c_void_p, c_double, c_void_p, c_void_p]),
    "pscf_ao_eval": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                             c_void_p]),
    "pscf_becke_weights": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                   c_void_p]),
    "pscf_b3lyp": (c_int, [c_int, c_void_p, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_b3lyp_host": (c_int, [c_void_p, c_int, c_double, c_void_p, c_void_p, c_void_p]),
    "pscf_xc_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "pscf_xc": (c_int, [c_int, c_void_p, c_int64, c_int64, c_int64, c_int, c_int, c_void_p, c_void_p, c_double,
                        c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "pscf_scf_state_bytes": (c_size_t, [c_int]),
    "pscf_scf_step": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "pscf_energy": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "pscf_cast": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int64, c_void_p]),
    # nuclear gradient
    "pscf_grad_int1e": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                c_void_p]),
    "pscf_grad_eri": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                              c_double, c_void_p, c_void_p]),
    "pscf_grad_xc_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "pscf_grad_xc": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                             c_double, c_void_p, c_size_t, c_void_p, c_void_p]),
    # batched form
    "pscf_batch_eri_s8": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_batch_eri_s8_div": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_batch_chunks": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64, c_void_p]),
    "pscf_batch_enuc": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_batch_int1e": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "pscf_rks_run": (c_int, [c_int, c_int, c_int, c_int, c_int, c_double, c_int, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p,
                             c_size_t, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_batch_eri_scale": (c_int, [c_int, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "pscf_batch_jk_s8": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double,
                                 c_void_p, c_int, c_int, c_int, c_void_p]),
    "pscf_batch_grid": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_batch_becke": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_batch_ao": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p]),
    "pscf_batch_xc": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_double, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64,
                              c_void_p]),
    "pscf_batch_xc_direct": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_double, c_void_p,
                                     c_void_p, c_int, c_int, c_void_p, c_void_p, c_int64, c_int64, c_void_p]),
    "pscf_batch_chol": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_batch_guess": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p]),
    "pscf_batch_scf_step": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_int, c_void_p]),
    "pscf_batch_energy": (c_int, [c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double, c_int, c_void_p]),
    # large molecules and the precision study
    "pscf_eri_s8_to_s1": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "pscf_jk_s1": (c_int, [c_int, c_void_p, c_int, c_void_p, c_double, c_void_p, c_void_p]),
    "pscf_jk_s8_large_workspace_bytes": (c_size_t, [c_int]),
    "pscf_jk_s8_large": (c_int, [c_int, c_void_p, c_int, c_void_p, c_double, c_void_p, c_void_p, c_size_t,
                                 c_void_p]),
    "pscf_gemm": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64,
                          c_void_p, c_int64, c_int, c_void_p]),
    "pscf_cast_to": (c_int, [c_int, c_void_p, c_void_p, c_int64, c_void_p]),
    "pscf_xc_large_parts": (c_int, [c_int]),
    "pscf_xc_large_stage": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "pscf_xc_large_finish": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "pscf_scf_step_ws_bytes": (c_size_t, [c_int]),
    "pscf_scf_step_ws": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_size_t,
                                 c_void_p]),
    # learned initial guess
    "pscf_batch_purify": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
    "pscf_pair_features": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_pair_targets": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_pair_assemble": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pscf_mlp_gemm": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "pscf_mlp_mse_grad": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    "pscf_mlp_colsum": (c_int, [c_int, c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "pscf_mlp_adam": (c_int, [c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                              c_float, c_int, c_void_p]),
    "pscf_mlp_to_bf16": (c_int, [c_int64, c_void_p, c_void_p, c_void_p]),
}

_LIB = None


class HIPError(RuntimeError):
    pass


def lib():
    """Load the library (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise HIPError(f"{LIB_PATH} not found: build it with `python -m pyscf_ipu_amd.build` "
                           "(the HIP path has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().pscf_last_error().decode(errors="replace")
        raise HIPError(f"{what or 'pscf call'} failed ({rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


def symbols():
    """Exported pscf_* symbol names the binding expects."""
    return list(_SIGS)
