"""ORACLE (test infrastructure only): ctypes front of cint_oracle.c.

Functions and what pins them (tests/test_oracle_integrals.py):
  boys        closed forms F_0 = sqrt(pi/T) erf(sqrt T)/2, F_m(0) = 1/(2m+1), mpmath
  rys         mpmath quadrature of sum_r w_r u_r^k = F_k(T) for k < 2n
  int1e       Szabo-Ostlund H2 STO-3G S/T/V (pyscf_ipu test/test_integrals.py:24-99),
              PyQuante primitive T = 1.5 and V = -1.595769
  eri_s8      Szabo-Ostlund H2 ERIs and PyQuante (ss|ss) = 1.128379, (ss|pp) = 0.940316
              (test_integrals.py:111-133); THO/MD second algorithm (md_eri.py)
  sparse_jk   dense einsum definition of J and K (nanoDFT.py:130-132)
  eri_ip1, int1e_ip   central finite differences of eri_s8 / int1e under a
              displaced atom (tests/test_grad_oracle.py)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile liboracle.so from cint_oracle.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def boys(mmax: int, T: float) -> np.ndarray:
    out = np.zeros(mmax + 1)
    lib().oracle_boys(ctypes.c_int(mmax), ctypes.c_double(T), _p(out))
    return out


def rys(n: int, T: float):
    u = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_rys(ctypes.c_int(n), ctypes.c_double(T), _p(u), _p(w))
    return u, w


def _mol_arrays(mol):
    atm = np.ascontiguousarray(mol._atm, dtype=np.int32)
    bas = np.ascontiguousarray(mol._bas, dtype=np.int32)
    env = np.ascontiguousarray(mol._env, dtype=np.float64)
    return atm, bas, env


def eri_s8(mol) -> np.ndarray:
    """Unique ERIs packed in PySCF's aosym='s8' order, FP64."""
    atm, bas, env = _mol_arrays(mol)
    N = mol.nao_nr()
    P = N * (N + 1) // 2
    out = np.zeros(P * (P + 1) // 2)
    rc = lib().oracle_eri_s8(_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)), _p(env), _p(out))
    if rc != 0:
        raise ValueError("oracle_eri_s8: unsupported shell")
    return out


def eri_quartet(mol, shls) -> np.ndarray:
    atm, bas, env = _mol_arrays(mol)
    dims = [2 * int(mol._bas[s, 1]) + 1 for s in shls]
    out = np.zeros(int(np.prod(dims)))
    sh = np.asarray(shls, dtype=np.int32)
    lib().oracle_eri_quartet(_p(atm), _p(bas), ctypes.c_int(len(bas)), _p(env), _p(sh), _p(out))
    return out.reshape(dims)


def int1e(mol):
    """(S, T, V) FP64 nao x nao."""
    atm, bas, env = _mol_arrays(mol)
    N = mol.nao_nr()
    S, T, V = np.zeros((N, N)), np.zeros((N, N)), np.zeros((N, N))
    rc = lib().oracle_int1e(_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)), _p(env),
                            ctypes.c_int(N), _p(S), _p(T), _p(V))
    if rc != 0:
        raise ValueError("oracle_int1e: unsupported shell")
    return S, T, V


def sparse_jk(eri: np.ndarray, idx: np.ndarray, dm: np.ndarray, hyb: float = 0.2) -> np.ndarray:
    """16-pass gather/scale/scatter J - hyb/2 K over multiplicity-divided ERIs."""
    N = dm.shape[0]
    eri = np.ascontiguousarray(eri, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int16).reshape(-1, 4)
    dm = np.ascontiguousarray(dm, dtype=np.float64)
    out = np.zeros((N, N))
    lib().oracle_sparse_jk(_p(eri), _p(idx), ctypes.c_int64(len(eri)), _p(dm), ctypes.c_int(N),
                           ctypes.c_double(hyb), _p(out))
    return out


def s8_to_s1(eri_s8: np.ndarray, N: int) -> np.ndarray:
    """Dense (N,N,N,N) tensor from the s8 packing (the role of reconstruct_ERI,
    pyscf_ipu/nanoDFT/compute_eri_utils.py:6-19)."""
    p, q = np.tril_indices(N)
    P = len(p)
    pair = np.zeros((N, N), dtype=np.int64)
    pair[p, q] = np.arange(P)
    pair[q, p] = np.arange(P)
    a, b = np.maximum(pair[:, :, None, None], pair[None, None, :, :]), np.minimum(pair[:, :, None, None], pair[None, None, :, :])
    return eri_s8[a * (a + 1) // 2 + b]


def eri_ip1(mol) -> np.ndarray:
    """(3, N, N, N, N) (nabla i j | k l), PySCF's int2e_ip1 layout (derivative
    on the electron coordinate of the first function).  Small molecules only."""
    atm, bas, env = _mol_arrays(mol)
    N = mol.nao_nr()
    out = np.zeros((3, N, N, N, N))
    rc = lib().oracle_eri_ip1(_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)), _p(env),
                              ctypes.c_int(N), _p(out))
    if rc != 0:
        raise ValueError("oracle_eri_ip1: unsupported shell")
    return out


def int1e_ip(mol, nuc: int = -1):
    """(ipovlp, ipkin, ipnuc, iprinv) each (3, N, N): <nabla i | O | j> for O =
    1, -nabla^2/2, sum_C -Z_C/|r - C|, and (nuc >= 0) 1/|r - R_nuc| (else None)."""
    atm, bas, env = _mol_arrays(mol)
    N = mol.nao_nr()
    o, k, v = np.zeros((3, N, N)), np.zeros((3, N, N)), np.zeros((3, N, N))
    r = np.zeros((3, N, N)) if nuc >= 0 else None
    rc = lib().oracle_int1e_ip(_p(atm), ctypes.c_int(len(atm)), _p(bas), ctypes.c_int(len(bas)), _p(env),
                               ctypes.c_int(N), _p(o), _p(k), _p(v), ctypes.c_int(nuc),
                               _p(r) if r is not None else None)
    if rc != 0:
        raise ValueError("oracle_int1e_ip: unsupported shell")
    return o, k, v, r
