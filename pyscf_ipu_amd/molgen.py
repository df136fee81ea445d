"""Seeded synthetic molecule generator for the benchmark workloads.

Heavy atoms are drawn from C/N/O with probabilities 0.7/0.15/0.15.  Each new
heavy atom bonds to a uniformly chosen earlier heavy atom that still has free
valence (C 4, N 3, O 2) at a bond length uniform in [1.30, 1.50] Angstrom along
a uniformly random direction; a placement closer than 1.1 Angstrom to any other
heavy atom is redrawn.  Hydrogens then fill every remaining valence at 1.09
Angstrom in random directions, a hydrogen closer than 0.9 Angstrom to any atom
being redrawn.  The result is a neutral closed-shell molecule with single bonds.
"""
from __future__ import annotations

from typing import List, Tuple

import numpy as np

VALENCE = {"C": 4, "N": 3, "O": 2}
HEAVY = ("C", "N", "O")
HEAVY_P = (0.7, 0.15, 0.15)


def _unit(rng: np.random.Generator) -> np.ndarray:
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def random_molecule(rng: np.random.Generator, n_heavy: int, max_tries: int = 10000) -> List[Tuple[str, np.ndarray]]:
    """One molecule as [(symbol, xyz Angstrom)] with `n_heavy` heavy atoms."""
    for _attempt in range(100):
        syms = list(rng.choice(HEAVY, size=n_heavy, p=HEAVY_P))
        pos = [np.zeros(3)]
        used = [0] * n_heavy
        ok = True
        for a in range(1, n_heavy):
            placed = False
            for _ in range(max_tries):
                free = [b for b in range(a) if used[b] < VALENCE[syms[b]]]
                if not free:
                    break
                parent = free[rng.integers(len(free))]
                r = rng.uniform(1.30, 1.50)
                cand = pos[parent] + r * _unit(rng)
                if all(np.linalg.norm(cand - p) >= 1.1 for p in pos):
                    pos.append(cand)
                    used[parent] += 1
                    used[a] += 1
                    placed = True
                    break
            if not placed:
                ok = False
                break
        if not ok:
            continue
        atoms = [(s, p) for s, p in zip(syms, pos)]
        all_pos = list(pos)
        for b in range(n_heavy):
            for _h in range(VALENCE[syms[b]] - used[b]):
                for _ in range(max_tries):
                    cand = pos[b] + 1.09 * _unit(rng)
                    if all(np.linalg.norm(cand - p) >= 0.9 for p in all_pos):
                        break
                else:
                    ok = False
                    break
                atoms.append(("H", cand))
                all_pos.append(cand)
            if not ok:
                break
        if ok:
            return atoms
    raise RuntimeError("molecule generator failed to place atoms")


def to_mol_str(atoms: List[Tuple[str, np.ndarray]]) -> str:
    return "; ".join(f"{s} {x:.8f} {y:.8f} {z:.8f}" for s, (x, y, z) in atoms)


def plain_set(n: int = 32, n_heavy: int = 9, seed: int = 0) -> List[str]:
    """`n` molecules with `n_heavy` heavy atoms each (9 heavy atoms: STO-3G nao 50-70)."""
    rng = np.random.default_rng(seed)
    return [to_mol_str(random_molecule(rng, n_heavy)) for _ in range(n)]


def batched_set(n: int = 4096, heavy_range=(1, 9), seed: int = 1) -> List[str]:
    """`n` molecules with heavy-atom count uniform over `heavy_range` (inclusive)."""
    rng = np.random.default_rng(seed)
    lo, hi = heavy_range
    return [to_mol_str(random_molecule(rng, int(rng.integers(lo, hi + 1)))) for _ in range(n)]
