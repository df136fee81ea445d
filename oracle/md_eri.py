"""ORACLE (test infrastructure only): second, independent ERI algorithm.

Restates the closed-form Taketa-Huzinaga-O-ohata (THO) primitive ERI of
pyscf_ipu/experimental/integrals.py:186-244 (`_eri_primitives`, helpers
`binom_factor` special.py:118, `gammanu` special.py:99) in numpy FP64, with the
Boys function from the C oracle.  Pure-Python loops: small molecules only.
Used to cross-check the Rys oracle (cint_oracle.c), which shares no code path
with it beyond the Boys function.
"""
from __future__ import annotations

import math
from itertools import product

import numpy as np

from . import cint


def _binom(n, k):
    return math.comb(n, k) if 0 <= k <= n else 0


def binom_factor(i, j, a, b, lmax):
    """f_s(i, j, a, b) = sum_t C(i, s-t) C(j, t) a^(i-s+t) b^(j-t) for s = 0..lmax."""
    out = np.zeros(lmax + 1)
    for s in range(lmax + 1):
        for t in range(s + 1):
            if s - i <= t <= j:
                out[s] += _binom(i, s - t) * _binom(j, t) * a ** (i - (s - t)) * b ** (j - t)
    return out


def _H(l1, l2, a, b, i, r, gamma):
    u = math.factorial(i) * binom_factor(l1, l2, a, b, l1 + l2)[i]
    v = math.factorial(r) * math.factorial(i - 2 * r) * (4 * gamma) ** (i - r)
    return u / v


def _c_term(la, lb, lc, ld, pa, pb, qc, qd, qp, g1, g2, delta):
    L = la + lb + lc + ld
    C = np.zeros(L + 1)
    for i1 in range(la + lb + 1):
        for i2 in range(lc + ld + 1):
            for r1 in range(i1 // 2 + 1):
                for r2 in range(i2 // 2 + 1):
                    for u in range((i1 + i2) // 2 - r1 - r2 + 1):
                        h = _H(la, lb, pa, pb, i1, r1, g1) * _H(lc, ld, qc, qd, i2, r2, g2)
                        index = i1 + i2 - 2 * (r1 + r2) - u
                        x = (-1) ** (i2 + u) * math.factorial(index + u) * qp ** (index - u)
                        y = math.factorial(u) * math.factorial(index - u) * delta ** index
                        C[index] += h * x / y
    return C


def eri_primitive(a, A, la, b, B, lb, c, C, lc, d, D, ld):
    """Unnormalised primitive ERI of cartesian Gaussians x^lx y^ly z^lz exp(-a r^2)."""
    A, B, C, D = map(np.asarray, (A, B, C, D))
    p, q = a + b, c + d
    P, Q = (a * A + b * B) / p, (c * C + d * D) / q
    pnorm = math.exp(-a * b / p * float(np.sum((A - B) ** 2)))
    qnorm = math.exp(-c * d / q * float(np.sum((C - D) ** 2)))
    delta = 1 / (4.0 * p) + 1 / (4.0 * q)
    Cs = [_c_term(la[x], lb[x], lc[x], ld[x], P[x] - A[x], P[x] - B[x], Q[x] - C[x], Q[x] - D[x], Q[x] - P[x], p, q, delta)
          for x in range(3)]
    T = float(np.sum((Q - P) ** 2)) / (4.0 * delta)
    nmax = sum(len(c_) - 1 for c_ in Cs)
    F = cint.boys(nmax, T)
    tot = 0.0
    for i, j, k in product(range(len(Cs[0])), range(len(Cs[1])), range(len(Cs[2]))):
        tot += Cs[0][i] * Cs[1][j] * Cs[2][k] * F[i + j + k]
    return 2.0 * math.pi ** 2 / (p * q) * math.sqrt(math.pi / (p + q)) * pnorm * qnorm * tot


_CART = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)]}


def eri_dense(mol) -> np.ndarray:
    """Dense (N,N,N,N) spherical ERIs of an s/p basis."""
    shells = mol.shell_arrays()
    funcs = []
    for (R, l, exps, coefs) in shells:
        for lmn in _CART[l]:
            funcs.append((R, lmn, exps, coefs))
    N = len(funcs)
    out = np.zeros((N, N, N, N))
    for i in range(N):
        for j in range(i + 1):
            for k in range(N):
                for l in range(k + 1):
                    if i * (i + 1) // 2 + j < k * (k + 1) // 2 + l:
                        continue
                    fa, fb, fc, fd = funcs[i], funcs[j], funcs[k], funcs[l]
                    v = 0.0
                    for ea, ca in zip(fa[2], fa[3]):
                        for eb, cb in zip(fb[2], fb[3]):
                            for ec, cc in zip(fc[2], fc[3]):
                                for ed, cd in zip(fd[2], fd[3]):
                                    v += ca * cb * cc * cd * eri_primitive(ea, fa[0], fa[1], eb, fb[0], fb[1],
                                                                           ec, fc[0], fc[1], ed, fd[0], fd[1])
                    for (a, b, c, d) in [(i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                                         (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)]:
                        out[a, b, c, d] = v
    return out
