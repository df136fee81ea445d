"""ORACLE (test infrastructure only): integer index arithmetic of the s8 path.

Restates, independently of the product's vectorised versions:
  get_i_j               pyscf_ipu/nanoDFT/sparse_symmetric_ERI.py:13-16
  num_repetitions_fast  sparse_symmetric_ERI.py:68-79
  index table           compute_indices.cpp:54-95 (== indices_func :81-84)
  shell quartet list    pyscf_ipu/electron_repulsion/direct.py:772-811
                        (filter i>=j, k>=l, i*j>=k*l; grouping by block size)
by brute-force enumeration in pure Python (small N only).  Bit-exact targets.
"""
from __future__ import annotations

import numpy as np


def tril_pairs(N: int):
    """[(i, j)] for i >= j in the order of the packed index c = i(i+1)/2 + j."""
    return [(i, j) for i in range(N) for j in range(i + 1)]


def get_i_j_bruteforce(vals, N: int):
    pairs = tril_pairs(N)
    i = np.array([pairs[v][0] for v in vals], dtype=np.int64)
    j = np.array([pairs[v][1] for v in vals], dtype=np.int64)
    return i, j


def multiplicity_bruteforce(i, j, k, l) -> int:
    """Number of distinct index tuples among the 8 permutation images."""
    imgs = {(i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
            (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)}
    return 8 // len(imgs)


def index_table(i, j, k, l, N):
    """The 32 flat indices of compute_indices.cpp:54-95 (J gather, J scatter, K gather, K scatter)."""
    return [i * N + j, j * N + i, i * N + j, j * N + i, k * N + l, l * N + k, k * N + l, l * N + k,
            k * N + l, k * N + l, l * N + k, l * N + k, i * N + j, i * N + j, j * N + i, j * N + i,
            k * N + j, k * N + i, l * N + j, l * N + i, i * N + l, i * N + k, j * N + l, j * N + k,
            i * N + l, j * N + l, i * N + k, j * N + k, k * N + j, l * N + j, k * N + i, l * N + i]


def shell_quartets_bruteforce(nbas: int):
    out = []
    for i in range(nbas):
        for j in range(nbas):
            for k in range(nbas):
                for l in range(nbas):
                    if i >= j and k >= l and i * j >= k * l:
                        out.append((i, j, k, l))
    return out


def einsum_inputs_bruteforce(bas: np.ndarray):
    """(indices (num_calls, 8) = di, dj, dk, dl, i0, j0, k0, l0 and do_lists
    (num_calls, 8)) of the shell-quartet loop of
    pyscf_ipu/electron_repulsion/direct.py:1029-1110: an index image of a
    block's origin is applied by the first quartet (in loop order) that
    produces it, so every s1 element is counted once."""
    nbas = bas.shape[0]
    ao_loc = np.cumsum(np.concatenate([[0], (bas[:, 1] * 2 + 1) * bas[:, 3]])).astype(int)
    seen = set()
    idx, dol = [], []
    for (i, j, k, l) in shell_quartets_bruteforce(nbas):
        d = [ao_loc[s + 1] - ao_loc[s] for s in (i, j, k, l)]
        i0, j0, k0, l0 = (ao_loc[s] for s in (i, j, k, l))
        images = [(i0, j0, k0, l0), (i0, j0, l0, k0), (j0, i0, k0, l0), (j0, i0, l0, k0),
                  (k0, l0, i0, j0), (k0, l0, j0, i0), (l0, k0, i0, j0), (l0, k0, j0, i0)]
        do = []
        for key in images:
            do.append(key not in seen)
            seen.add(key)
        idx.append(d + [i0, j0, k0, l0])
        dol.append(do)
    return np.array(idx, dtype=np.int32), np.array(dol, dtype=np.int32)
