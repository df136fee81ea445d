"""Becke-partitioned molecular integration grids (host set-up, FP64).

The reference builds its grid with PySCF (`pyscf.dft.gen_grid.Grids(mol)`,
`grids.level = opts.level`, pyscf_ipu/nanoDFT/nanoDFT.py:215-217) and carries a
local copy of the partition step (pyscf_ipu/pyscf_utils/build_grid.py:34-134).
This module restates that scheme:

* radial: Treutler-Ahlrichs M4 mapping (alpha = 0.6) of Chebyshev nodes,
  r = -ln(2)^-1 (1+x)^0.6 ln((1-x)/2);
* angular: Lebedev-Laikov octahedral rules.  The point sets are *derived* here
  by solving the exactness equations of each rule's orbit structure with
  Newton's method (no tables are copied); tests check exactness on spherical
  harmonics;
* per-element sizes by grid level (radial count, Lebedev order) and NWChem
  pruning of the angular order by r / r_Bragg;
* Becke's fuzzy-cell partition (three smoothing iterations,
  build_grid.py:34-39) with Treutler's Bragg-radius adjustment
  (build_grid.py:44-100 passes `radii_adjust`).

Parity against PySCF's own point set is unpinned here (no PySCF in this
environment); tests pin the pieces by quadrature exactness and by
int rho = N_electrons.  Both the GPU path and the oracle consume the same grid,
so it is an input of the hot path, not part of it.  The per-point Becke
weights are also produced on the GPU (`becke_weights` entry point); this
module's FP64 version is the host set-up path and the reference for it.
"""
from __future__ import annotations

import functools
from fractions import Fraction
from math import comb, factorial
from typing import Dict, List, Tuple

import numpy as np

from .gto import BOHR

# --------------------------------------------------------------------- sizes ---
# Radial points / Lebedev order by grid level (rows) and period (cols 1..3).
RAD_GRIDS = np.array([(10, 15, 20), (30, 40, 50), (40, 60, 65), (50, 75, 80), (60, 90, 95),
                      (70, 105, 110), (80, 120, 125), (90, 135, 140), (100, 150, 155), (200, 200, 200)])
ANG_ORDER = np.array([(11, 13, 13), (15, 17, 17), (17, 19, 19), (19, 23, 23), (29, 29, 29),
                      (35, 41, 41), (41, 47, 47), (47, 53, 53), (59, 59, 59), (131, 131, 131)])
LEBEDEV_ORDER_TO_N = {3: 6, 5: 14, 7: 26, 9: 38, 11: 50, 13: 74, 15: 86, 17: 110, 19: 146, 21: 170, 23: 194}
LEBEDEV_NGRID = np.array([1, 6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194, 230, 266, 302, 350, 434])

# Bragg-Slater radii (Angstrom) for H..Ne, index = nuclear charge.
_BRAGG_A = np.array([0.0, 0.35, 1.40, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 1.50])
BRAGG_RADII = _BRAGG_A / BOHR


def _period(z: int) -> int:
    return 0 if z <= 2 else (1 if z <= 10 else 2)


# ------------------------------------------------------------------ Lebedev ---
# Orbit types of the octahedral group: a1 (6), a2 (12), a3 (8), b(l) (24),
# c(p) (24), d(r, s) (48).  Each rule is a set of orbits with unknown weights
# (and positions for b/c/d) fixed by exact integration of the invariant
# polynomials p2^a p3^b, p2 = x2y2 + y2z2 + z2x2, p3 = x2y2z2, of degree <= L.
_RULES = {
    # npts: (order, orbit types, starting guess for Newton: weights..., params...)
    6: (3, ["a1"], [1 / 6], []),
    14: (5, ["a1", "a3"], [0.0666667, 0.075], []),
    26: (7, ["a1", "a2", "a3"], [0.0476190, 0.0380952, 0.0321429], []),
    38: (9, ["a1", "a3", "c"], [0.0095238, 0.0321429, 0.0285714], [0.4597008]),
    50: (11, ["a1", "a2", "a3", "b"], [0.0126984, 0.0225749, 0.0210937, 0.0201733], [0.3015113]),
    74: (13, ["a1", "a2", "a3", "b", "c"], [0.000513, 0.0166040, -0.0295860, 0.0265762, 0.0165222],
         [0.4803845, 0.3207726]),
    86: (15, ["a1", "a3", "b", "b", "c"], [0.0115440, 0.0119439, 0.0111106, 0.0118765, 0.0118123],
         [0.3696028, 0.6943540, 0.3742430]),
    110: (17, ["a1", "a3", "b", "b", "b", "c"], [0.0038283, 0.0097937, 0.0082117, 0.0099428, 0.0095955, 0.0096950],
          [0.1851156, 0.6904210, 0.3956895, 0.4783690]),
}


def _orbit_points(kind: str, params: List[float]) -> np.ndarray:
    pts = []
    if kind == "a1":
        base = [(1.0, 0.0, 0.0)]
    elif kind == "a2":
        s = np.sqrt(0.5)
        base = [(0.0, s, s)]
    elif kind == "a3":
        s = np.sqrt(1 / 3)
        base = [(s, s, s)]
    elif kind == "b":
        l = params[0]
        m = np.sqrt(1 - 2 * l * l)
        base = [(l, l, m)]
    elif kind == "c":
        p = params[0]
        q = np.sqrt(1 - p * p)
        base = [(p, q, 0.0)]
    elif kind == "d":
        r, s = params
        t = np.sqrt(1 - r * r - s * s)
        base = [(r, s, t)]
    else:
        raise ValueError(kind)
    for b in base:
        for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
            v = np.array([b[perm[0]], b[perm[1]], b[perm[2]]])
            for signs in np.ndindex(2, 2, 2):
                pts.append(v * (1 - 2 * np.array(signs)))
    pts = np.unique(np.round(np.array(pts), 15), axis=0)
    return pts


_NPARAM = {"a1": 0, "a2": 0, "a3": 0, "b": 1, "c": 1, "d": 2}


def _sphere_monomial_avg(a: int, b: int, c: int) -> Fraction:
    """<x^2a y^2b z^2c> over the unit sphere = (2a-1)!!(2b-1)!!(2c-1)!!/(2(a+b+c)+1)!!."""
    def df(n):
        r = 1
        while n > 1:
            r *= n
            n -= 2
        return r
    return Fraction(df(2 * a - 1) * df(2 * b - 1) * df(2 * c - 1), df(2 * (a + b + c) + 1))


def _invariant_exponents(order: int) -> List[Tuple[int, int]]:
    return [(a, b) for b in range(order // 6 + 1) for a in range(order // 4 + 1) if 4 * a + 6 * b <= order]


def _invariant_exact(a: int, b: int) -> float:
    """Sphere average of p2^a p3^b via the multinomial expansion of p2^a."""
    tot = Fraction(0)
    for i in range(a + 1):
        for j in range(a - i + 1):
            k = a - i - j
            coef = factorial(a) // (factorial(i) * factorial(j) * factorial(k))
            # (x2y2)^i (y2z2)^j (z2x2)^k (x2y2z2)^b
            tot += coef * _sphere_monomial_avg(i + k + b, i + j + b, j + k + b)
    return float(tot)


def _invariant_values(pts: np.ndarray, exps) -> np.ndarray:
    x2, y2, z2 = pts[:, 0] ** 2, pts[:, 1] ** 2, pts[:, 2] ** 2
    p2 = x2 * y2 + y2 * z2 + z2 * x2
    p3 = x2 * y2 * z2
    return np.stack([p2 ** a * p3 ** b for a, b in exps])


def _rule_residual(kinds, order, x):
    nw = len(kinds)
    w, params = x[:nw], x[nw:]
    exps = _invariant_exponents(order)
    exact = np.array([_invariant_exact(a, b) for a, b in exps])
    tot = np.zeros(len(exps))
    pi = 0
    for kind, wk in zip(kinds, w):
        npar = _NPARAM[kind]
        pts = _orbit_points(kind, list(params[pi:pi + npar]))
        pi += npar
        tot += wk * _invariant_values(pts, exps).sum(axis=1)
    return tot - exact


@functools.lru_cache(maxsize=None)
def lebedev(npts: int) -> np.ndarray:
    """(npts, 4) array of unit vectors and weights summing to 1."""
    order, kinds, w0, p0 = _RULES[npts]
    x = np.array(w0 + p0, dtype=np.float64)
    for _ in range(100):
        f = _rule_residual(kinds, order, x)
        if np.max(np.abs(f)) < 1e-16:
            break
        J = np.empty((len(f), len(x)))
        for k in range(len(x)):
            h = 1e-7 * max(1.0, abs(x[k]))
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            J[:, k] = (_rule_residual(kinds, order, xp) - _rule_residual(kinds, order, xm)) / (2 * h)
        dx = np.linalg.lstsq(J, -f, rcond=None)[0]
        x = x + dx
        if np.max(np.abs(dx)) < 1e-15:
            break
    f = _rule_residual(kinds, order, x)
    if np.max(np.abs(f)) > 1e-13:
        raise RuntimeError(f"Lebedev {npts}: exactness residual {np.max(np.abs(f))}")
    nw = len(kinds)
    out = []
    pi = 0
    for kind, wk in zip(kinds, x[:nw]):
        npar = _NPARAM[kind]
        pts = _orbit_points(kind, list(x[nw + pi:nw + pi + npar]))
        pi += npar
        out.append(np.hstack([pts, np.full((len(pts), 1), wk)]))
    grid = np.vstack(out)
    assert grid.shape[0] == npts, (npts, grid.shape)
    return grid


# ------------------------------------------------------------------- radial ---
def treutler_ahlrichs(n: int) -> Tuple[np.ndarray, np.ndarray]:
    """Treutler-Ahlrichs M4 radial nodes and Jacobians dr (ascending r)."""
    i = np.arange(1, n + 1)
    step = np.pi / (n + 1)
    x = np.cos(i * step)
    ln2 = 1.0 / np.log(2.0)
    r = -ln2 * (1 + x) ** 0.6 * np.log((1 - x) / 2)
    dr = step * np.sin(i * step) * ln2 * (1 + x) ** 0.6 * (-0.6 / (1 + x) * np.log((1 - x) / 2) + 1 / (1 - x))
    return r[::-1].copy(), dr[::-1].copy()


def nwchem_prune(z: int, rads: np.ndarray, n_ang: int) -> np.ndarray:
    """Angular point count per radial shell, NWChem scheme keyed on r / r_Bragg."""
    alphas = np.array([(0.25, 0.5, 1.0, 4.5), (0.1667, 0.5, 0.9, 3.5), (0.1, 0.4, 0.8, 2.5)])
    leb = LEBEDEV_NGRID[4:]
    if n_ang < 50:
        return np.full(len(rads), n_ang)
    if n_ang == 50:
        leb_l = np.array([1, 2, 2, 2, 1])
    else:
        idx = int(np.where(leb == n_ang)[0][0])
        leb_l = np.array([1, 3, idx - 1, idx, idx - 1])
    r_atom = BRAGG_RADII[z] + 1e-200
    place = ((rads / r_atom).reshape(-1, 1) > alphas[_period(z)]).sum(axis=1)
    return leb[leb_l[place]]


@functools.lru_cache(maxsize=None)
def atomic_grid(z: int, level: int) -> Tuple[np.ndarray, np.ndarray]:
    """Atom-centred grid (coords relative to the nucleus, Bohr) and weights
    (4 pi r^2 dr x Lebedev weight), before Becke partitioning."""
    p = _period(z)
    n_rad = int(RAD_GRIDS[level, p])
    n_ang = LEBEDEV_ORDER_TO_N[int(ANG_ORDER[level, p])]
    rad, dr = treutler_ahlrichs(n_rad)
    rw = 4 * np.pi * rad ** 2 * dr
    angs = nwchem_prune(z, rad, n_ang)
    coords, vols = [], []
    for n in sorted(set(angs.tolist())):
        g = lebedev(int(n))
        sel = np.where(angs == n)[0]
        coords.append((rad[sel, None, None] * g[None, :, :3]).reshape(-1, 3))
        vols.append((rw[sel, None] * g[None, :, 3]).reshape(-1))
    return np.vstack(coords), np.concatenate(vols)


# ------------------------------------------------------------------- Becke ----
def becke_adjust_table(charges: np.ndarray) -> np.ndarray:
    """Treutler's size adjustment a_ij from sqrt(Bragg radius) ratios, clipped to +-0.5."""
    rad = np.sqrt(BRAGG_RADII[charges]) + 1e-200
    rr = rad.reshape(-1, 1) * (1.0 / rad)
    a = 0.25 * (rr.T - rr)
    return np.clip(a, -0.5, 0.5)


def becke_partition(coords: np.ndarray, atom_coords: np.ndarray, charges: np.ndarray, owner: np.ndarray,
                    chunk: int = 4096) -> np.ndarray:
    """Becke cell weight of each point for its owning atom: P_owner / sum_a P_a."""
    natm = len(atom_coords)
    a = becke_adjust_table(charges)
    rab = np.linalg.norm(atom_coords[:, None] - atom_coords[None], axis=-1)
    out = np.empty(len(coords))
    for s in range(0, len(coords), chunk):
        c = coords[s:s + chunk]
        d = np.linalg.norm(c[None] - atom_coords[:, None], axis=-1)  # (natm, g)
        pb = np.ones_like(d)
        for i in range(natm):
            for j in range(i):
                g = (d[i] - d[j]) / rab[i, j]
                g = g + a[i, j] * (1 - g * g)
                for _ in range(3):
                    g = (3 - g * g) * g * 0.5
                pb[i] *= 0.5 * (1 - g)
                pb[j] *= 0.5 * (1 + g)
        o = owner[s:s + chunk]
        out[s:s + chunk] = pb[o, np.arange(len(o))] / pb.sum(axis=0)
    return out


class Grid:
    """Molecular grid: `coords` (g,3) Bohr, `weights` (g,), and per-point owner
    atom and un-partitioned volume (`vol`), which the GPU Becke kernel uses."""

    def __init__(self, coords, weights, owner, vol):
        self.coords, self.weights, self.owner, self.vol = coords, weights, owner, vol

    @property
    def size(self) -> int:
        return len(self.coords)


def atomic_grids(mol, level: int = 0) -> Grid:
    """Concatenated atom-centred grids (weights not yet partitioned)."""
    ac = mol.atom_coords()
    z = mol.atom_charges()
    coords, vol, owner = [], [], []
    for ia in range(mol.natm):
        c, v = atomic_grid(int(z[ia]), level)
        coords.append(c + ac[ia])
        vol.append(v)
        owner.append(np.full(len(v), ia, dtype=np.int32))
    coords, vol, owner = np.vstack(coords), np.concatenate(vol), np.concatenate(owner)
    return Grid(coords, None, owner, vol)


def build_grid(mol, level: int = 0) -> Grid:
    """Full host build: atom grids then Becke partition, FP64."""
    g = atomic_grids(mol, level)
    g.weights = g.vol * becke_partition(g.coords, mol.atom_coords(), mol.atom_charges(), g.owner)
    return g
