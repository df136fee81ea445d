"""Algorithmic work per launch of each kernel group (frozen models for rooflines).

eri_s8   FP VALU-bound.  FLOPs = sum over primitive quartets of F(class), with
         F = 40 (ket pair + prefactor) + 2*(TAYLOR+1) + 4*(2R-1) + 1 (Boys)
           + ROOTS[R] (Rys roots from moments)
           + R * (10 + 3*(8 + 5*(n+1)(m+1) + 3*prod(l+1)) + 4*nc)
         (per root: B coefficients, per axis C coefficients + 2D recurrence over
         (n, m) = (li+lj, lk+ll) + horizontal transfer, then 4 flops per output
         component for w*Ix*Iy*Iz accumulation), R = (li+lj+lk+ll)/2 + 1.
jk_s8    HBM-bound.  Bytes = (#unique s8 ERIs) * sizeof(dtype): each ERI is read
         once; dm and the J/K accumulator live in LDS, indices are regenerated.
xc       HBM-bound.  Bytes = 4 g N sizeof(dtype) (AO values + gradients) +
         g sizeof(dtype) (weights).
scf_step FP64, latency-bound.  FLOPs = 38 N^3 (DIIS products 6N^3, L H L^T 4N^3,
         warm-start transform 4N^3, three Jacobi sweeps 18N^3, C and D 4N^3,
         extrapolation and errors 2N^3).
"""
from __future__ import annotations

import numpy as np

TAYLOR = 8
ROOTS = {1: 1, 2: 25, 3: 120}


def eri_flops_per_primitive(li, lj, lk, ll) -> int:
    R = (li + lj + lk + ll) // 2 + 1
    n, m = li + lj, lk + ll
    nc = (2 * li + 1) * (2 * lj + 1) * (2 * lk + 1) * (2 * ll + 1)
    per_root = 10 + 3 * (8 + 5 * (n + 1) * (m + 1) + 3 * (li + 1) * (lj + 1) * (lk + 1) * (ll + 1)) + 4 * nc
    return 40 + 2 * (TAYLOR + 1) + 4 * (2 * R - 1) + 1 + ROOTS[R] + R * per_root


def eri_s8_flops(mol) -> float:
    """Algorithmic FLOPs of one pscf_eri_s8 call (all unique shell quartets)."""
    return _eri_flops_shells(mol._bas[:, 1].astype(int), mol._bas[:, 2].astype(int))


def batch_eri_flops(b) -> float:
    """Algorithmic FLOPs of pscf_batch_eri_s8 over every molecule of a
    BatchRKS (shells bas0 .. bas0 + nbas of each molecule)."""
    tot = 0.0
    for m in range(b.M):
        b0, nb = int(b.desc[m, 5]), int(b.desc[m, 1])
        tot += _eri_flops_shells(b.bas[b0:b0 + nb, 1].astype(int), b.bas[b0:b0 + nb, 2].astype(int))
    return tot


def _eri_flops_shells(l, npr) -> float:
    pairs = [(i, j) for i in range(len(l)) for j in range(i + 1)]
    # aggregate pairs by (l_a, l_b sorted, nprim product)
    from collections import Counter
    pc = Counter()
    for i, j in pairs:
        la, lb = max(l[i], l[j]), min(l[i], l[j])
        pc[(la, lb, npr[i] * npr[j])] += 1
    keys = list(pc)
    tot = 0.0
    for a in range(len(keys)):
        for b in range(a + 1):
            ka, kb = keys[a], keys[b]
            cnt = pc[ka] * pc[kb] if a != b else pc[ka] * (pc[ka] + 1) // 2
            bra, ket = (ka, kb) if ka[0] + ka[1] >= kb[0] + kb[1] else (kb, ka)
            f = eri_flops_per_primitive(bra[0], bra[1], ket[0], ket[1])
            tot += cnt * ka[2] * kb[2] * f
    return tot


def per_launch_work(kernel: str, mols, dtype: str) -> dict:
    es = 8 if dtype == "fp64" else 4
    if kernel == "eri_s8":
        f = float(np.mean([eri_s8_flops(m) for m in mols]))
        return {"bound": "valu", "flops": f, "model": "Rys ERI op-count model (pyscf_ipu_amd/roofline.py), "
                                                     "FP VALU; peak = MI355X FP64/FP32 vector rate"}
    if kernel == "jk_s8":
        b = float(np.mean([(m.nao_nr() * (m.nao_nr() + 1) // 2) * (m.nao_nr() * (m.nao_nr() + 1) // 2 + 1) // 2
                           for m in mols])) * es
        return {"bound": "hbm", "bytes": b, "model": "sizeof(dtype) per unique s8 ERI"}
    if kernel == "xc":
        from pyscf_ipu_amd import grids
        b = float(np.mean([(4 * m.nao_nr() + 1) * grids.atomic_grids(m, 0).size for m in mols])) * es
        return {"bound": "hbm", "bytes": b, "model": "4 g N + g elements (AO values/gradients, weights)"}
    if kernel == "scf_step":
        f = float(np.mean([38.0 * m.nao_nr() ** 3 for m in mols]))
        # FP64 arithmetic (MFMA products plus VALU rotations) whose peak is the
        # MI355X FP64 rate; what limits it is latency: one workgroup per molecule
        return {"bound": "mfma", "flops": f, "model": "38 N^3 FP64",
                "limiter": "latency: one workgroup (one CU) per molecule; its eigen-decomposition is a chain of "
                            "dependent reflector, bisection or refinement steps"}
    raise KeyError(kernel)
