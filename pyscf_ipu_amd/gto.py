"""Molecule and basis description in libcint's atm/bas/env layout.

The reference builds molecules with PySCF (`pyscf_ipu/nanoDFT/nanoDFT.py:543`,
`build_mol`) and hands `mol._atm, mol._bas, mol._env` to its ERI driver
(`pyscf_ipu/electron_repulsion/direct.py:710-739`, `prepare_int_floats`).  PySCF
is not a dependency here, so this module provides a duck-typed `Mole` with the
same attributes and the handful of methods the SCF path calls
(`nao_nr`, `nelectron`, `energy_nuc`, `atom_coords`, `atom_charges`).

Layout (libcint conventions, kept so the integer/float input arrays of
`prepare_electron_repulsion_integrals` are byte-compatible):

* `_atm[ia] = [charge, ptr_coord, nuc_mod, ptr_zeta, ptr_frac_charge, 0]`
* `_bas[ib] = [atom, l, nprim, nctr, kappa, ptr_exp, ptr_coeff, 0]`
* `_env[:20]` reserved, then per-atom coordinates (Bohr) and per-shell
  exponents / radially normalised contraction coefficients.

AO order per atom is all s shells then all p shells; each p shell is
(px, py, pz).  Each contracted AO has unit norm.
"""
from __future__ import annotations

import math
import re
from typing import Dict, List, Sequence, Tuple

import numpy as np

BOHR = 0.52917721092  # Angstrom per Bohr (value PySCF 2.2 uses)
PTR_ENV_START = 20
ATOM_SLOTS = 6
BAS_SLOTS = 8
CHARGE_OF, PTR_COORD = 0, 1
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = 0, 1, 2, 3, 4, 5, 6

# Real-solid-harmonic angular factors folded in by libcint's cart->sph step for
# l = 0 and l = 1: 1/sqrt(4 pi) and sqrt(3/(4 pi)).
ANGULAR_FACTOR = (0.282094791773878143, 0.488602511902919921)

ELEMENTS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
SYMBOLS = {v: k for k, v in ELEMENTS.items()}

# --- basis sets -------------------------------------------------------------
# Published Gaussian basis sets (Hehre-Stewart-Pople STO-3G; Pople 6-31G), as
# distributed in the EMSL/Basis Set Exchange library.  Each entry is a list of
# (label, [(exponent, coef_s[, coef_p]), ...]); "SP" shells share exponents.
_STO3G = {
    "H": [("S", [(3.42525091, 0.15432897), (0.62391373, 0.53532814), (0.16885540, 0.44463454)])],
    "C": [("S", [(71.6168370, 0.15432897), (13.0450960, 0.53532814), (3.5305122, 0.44463454)]),
          ("SP", [(2.9412494, -0.09996723, 0.15591627), (0.6834831, 0.39951283, 0.60768372),
                  (0.2222899, 0.70011547, 0.39195739)])],
    "N": [("S", [(99.1061690, 0.15432897), (18.0523120, 0.53532814), (4.8856602, 0.44463454)]),
          ("SP", [(3.7804559, -0.09996723, 0.15591627), (0.8784966, 0.39951283, 0.60768372),
                  (0.2857144, 0.70011547, 0.39195739)])],
    "O": [("S", [(130.7093200, 0.15432897), (23.8088610, 0.53532814), (6.4436083, 0.44463454)]),
          ("SP", [(5.0331513, -0.09996723, 0.15591627), (1.1695961, 0.39951283, 0.60768372),
                  (0.3803890, 0.70011547, 0.39195739)])],
}

_631G = {
    "H": [("S", [(18.7311370, 0.03349460), (2.8253937, 0.23472695), (0.6401217, 0.81375733)]),
          ("S", [(0.1612778, 1.0)])],
    "C": [("S", [(3047.5249000, 0.0018347), (457.3695100, 0.0140373), (103.9486900, 0.0688426),
                 (29.2101550, 0.2321844), (9.2866630, 0.4679413), (3.1639270, 0.3623120)]),
          ("SP", [(7.8682724, -0.1193324, 0.0689991), (1.8812885, -0.1608542, 0.3164240),
                  (0.5442493, 1.1434564, 0.7443083)]),
          ("SP", [(0.1687144, 1.0, 1.0)])],
    "N": [("S", [(4173.5110000, 0.0018348), (627.4579000, 0.0139950), (142.9021000, 0.0685870),
                 (40.2343300, 0.2322410), (12.8202100, 0.4690700), (4.3904370, 0.3604550)]),
          ("SP", [(11.6263580, -0.1149610, 0.0675800), (2.7162800, -0.1691180, 0.3239070),
                  (0.7722180, 1.1458520, 0.7408950)]),
          ("SP", [(0.2120313, 1.0, 1.0)])],
    "O": [("S", [(5484.6717000, 0.0018311), (825.2349500, 0.0139501), (188.0469600, 0.0684451),
                 (52.9645000, 0.2327143), (16.8975700, 0.4701930), (5.7996353, 0.3585209)]),
          ("SP", [(15.5396160, -0.1107775, 0.0708743), (3.5999336, -0.1480263, 0.3397528),
                  (1.0137618, 1.1307670, 0.7271586)]),
          ("SP", [(0.2700058, 1.0, 1.0)])],
}

BASIS_SETS = {"sto3g": _STO3G, "631g": _631G}


def _basis_key(name: str) -> str:
    key = name.lower().replace("-", "").replace("_", "")
    if key not in BASIS_SETS:
        raise KeyError(f"basis {name!r} not available (have sto-3g, 6-31g)")
    return key


def element_shells(symbol: str, basis: str) -> List[Tuple[int, np.ndarray, np.ndarray]]:
    """Shells of one element as (l, exponents, raw coefficients), s shells first.

    SP shells are split into one s and one p shell with shared exponents, and the
    shells are ordered by angular momentum as PySCF's basis parser does.
    """
    table = BASIS_SETS[_basis_key(basis)]
    if symbol not in table:
        raise KeyError(f"element {symbol} not in basis {basis}")
    by_l: Dict[int, list] = {0: [], 1: []}
    for label, prims in table[symbol]:
        exps = np.array([p[0] for p in prims], dtype=np.float64)
        if label == "S":
            by_l[0].append((0, exps, np.array([p[1] for p in prims])))
        elif label == "SP":
            by_l[0].append((0, exps, np.array([p[1] for p in prims])))
            by_l[1].append((1, exps, np.array([p[2] for p in prims])))
        else:
            raise ValueError(label)
    return by_l[0] + by_l[1]


def gaussian_int(n: float, alpha: np.ndarray) -> np.ndarray:
    """int_0^inf x^n exp(-alpha x^2) dx."""
    n1 = (n + 1) * 0.5
    return math.gamma(n1) / (2.0 * alpha ** n1)


def normalized_coefficients(l: int, exps: np.ndarray, coefs: np.ndarray) -> np.ndarray:
    """Radial normalisation of a contracted shell, as stored in libcint's env.

    Each primitive gets 1/sqrt(int r^(2l+2) exp(-2 a r^2) dr), then the contraction
    is rescaled so that int R(r)^2 r^2 dr = 1.
    """
    cs = coefs * (1.0 / np.sqrt(gaussian_int(2 * l + 2, 2.0 * exps)))
    ee = gaussian_int(2 * l + 2, exps[:, None] + exps[None, :])
    s = float(cs @ ee @ cs)
    return cs / math.sqrt(s)


# --- molecule -----------------------------------------------------------------
def parse_atom_string(mol_str: str) -> List[Tuple[str, np.ndarray]]:
    """Parse "C 0 0 0; H 0.6 0.6 0.6" (Angstrom) into [(symbol, xyz)]."""
    atoms = []
    for chunk in re.split(r"[;\n]", mol_str):
        chunk = chunk.strip()
        if not chunk:
            continue
        parts = chunk.replace(",", " ").split()
        if len(parts) != 4:
            raise ValueError(f"cannot parse atom entry {chunk!r}")
        sym = parts[0].capitalize()
        if sym not in ELEMENTS:
            raise ValueError(f"unknown element {parts[0]!r}")
        atoms.append((sym, np.array([float(x) for x in parts[1:]], dtype=np.float64)))
    if not atoms:
        raise ValueError("empty molecule string")
    return atoms


class Mole:
    """Duck-typed stand-in for `pyscf.gto.Mole` restricted to s/p basis sets.

    Mirrors the attributes the reference's hot path reads: `_atm, _bas, _env`,
    `nao_nr()`, `nelectron`, `energy_nuc()`, `atom_coords()`, `basis`, `cart`.
    """

    cart = False

    def __init__(self, atom, basis: str = "sto-3g", unit: str = "Angstrom", charge: int = 0, spin: int = 0):
        if spin != 0:
            raise ValueError("restricted closed-shell only (spin=0)")
        if isinstance(atom, str):
            atoms = parse_atom_string(atom)
        else:
            atoms = [(s if isinstance(s, str) else SYMBOLS[int(s)], np.asarray(x, dtype=np.float64)) for s, x in atom]
        scale = 1.0 / BOHR if unit.lower().startswith("a") else 1.0
        self.basis = basis
        self.unit = unit
        self.charge = charge
        self.spin = spin
        self.elements = [s for s, _ in atoms]
        coords = np.array([x for _, x in atoms], dtype=np.float64).reshape(-1, 3) * scale
        self._build(coords)

    def _build(self, coords: np.ndarray) -> None:
        natm = len(self.elements)
        env: List[float] = [0.0] * PTR_ENV_START
        atm = np.zeros((natm, ATOM_SLOTS), dtype=np.int32)
        for ia, sym in enumerate(self.elements):
            atm[ia, CHARGE_OF] = ELEMENTS[sym]
            atm[ia, PTR_COORD] = len(env)
            env.extend(coords[ia].tolist())
            atm[ia, 2] = 1  # point nucleus
            atm[ia, 3] = len(env)
            env.append(0.0)
        bas = []
        shell_cache: Dict[str, List[Tuple[int, int, int]]] = {}
        for ia, sym in enumerate(self.elements):
            if sym not in shell_cache:
                entries = []
                for l, exps, coefs in element_shells(sym, self.basis):
                    pexp = len(env)
                    env.extend(exps.tolist())
                    pcoef = len(env)
                    env.extend(normalized_coefficients(l, exps, coefs).tolist())
                    entries.append((l, len(exps), pexp, pcoef))
                shell_cache[sym] = entries
            for l, nprim, pexp, pcoef in shell_cache[sym]:
                bas.append([ia, l, nprim, 1, 0, pexp, pcoef, 0])
        self._atm = atm
        self._bas = np.array(bas, dtype=np.int32).reshape(-1, BAS_SLOTS)
        self._env = np.array(env, dtype=np.float64)
        self.natm = natm
        self.nbas = len(bas)

    # -- PySCF-compatible queries --------------------------------------------
    def atom_coords(self) -> np.ndarray:
        """Nuclear coordinates in Bohr, (natm, 3)."""
        ptr = self._atm[:, PTR_COORD]
        return np.stack([self._env[p:p + 3] for p in ptr])

    def atom_charges(self) -> np.ndarray:
        return self._atm[:, CHARGE_OF].astype(np.int64)

    def atom_symbol(self, ia: int) -> str:
        return self.elements[ia]

    @property
    def nelectron(self) -> int:
        return int(self.atom_charges().sum()) - self.charge

    def ao_loc_nr(self) -> np.ndarray:
        dims = (self._bas[:, ANG_OF] * 2 + 1) * self._bas[:, NCTR_OF]
        return np.concatenate([[0], np.cumsum(dims)]).astype(np.int32)

    def nao_nr(self) -> int:
        return int(self.ao_loc_nr()[-1])

    def energy_nuc(self) -> float:
        z = self.atom_charges().astype(np.float64)
        r = self.atom_coords()
        i, j = np.tril_indices(self.natm, -1)
        d = np.sqrt(((r[i] - r[j]) ** 2).sum(axis=1))
        return float(np.sum(z[i] * z[j] / d))

    def shell_arrays(self):
        """Per-shell (centre, l, exps, coefs incl. the angular factor) lists."""
        coords = self.atom_coords()
        out = []
        for ib in range(self.nbas):
            ia, l, nprim = self._bas[ib, ATOM_OF], self._bas[ib, ANG_OF], self._bas[ib, NPRIM_OF]
            pe, pc = self._bas[ib, PTR_EXP], self._bas[ib, PTR_COEFF]
            out.append((coords[ia].copy(), int(l), self._env[pe:pe + nprim].copy(),
                        self._env[pc:pc + nprim].copy() * ANGULAR_FACTOR[l]))
        return out

    def mol_str(self) -> str:
        r = self.atom_coords() * BOHR
        return "; ".join(f"{s} {x:.10f} {y:.10f} {z:.10f}" for s, (x, y, z) in zip(self.elements, r))

    def __repr__(self) -> str:
        return f"Mole(natm={self.natm}, nao={self.nao_nr()}, nelec={self.nelectron}, basis={self.basis})"


def build_mol(mol_str: str, basis_name: str = "sto-3g") -> Mole:
    """Counterpart of `build_mol` (`pyscf_ipu/nanoDFT/nanoDFT.py:543-546`)."""
    return Mole(mol_str, basis=basis_name, unit="Angstrom")


# Methane geometry of the plain-method CPU case (C at the origin, H on the
# tetrahedron vertices at 0.6276 Angstrom per axis).
METHANE = ("C 0 0 0; H 0.6276 0.6276 0.6276; H 0.6276 -0.6276 -0.6276; "
           "H -0.6276 0.6276 -0.6276; H -0.6276 -0.6276 0.6276")
