"""QM1B-schema dataset rows from converged SCF results.

The reference's dataset (qm1b/README.md:7-18) is Apache Parquet with one row
per conformer:

  smile   SMILES string of the molecule (several conformers share one)
  atoms   atom symbols concatenated, e.g. "CCOHHHH" (the order of `pos`)
  z       atomic numbers (int), the same order
  energy  total energy (eV)
  homo    HOMO orbital energy (eV)
  lumo    LUMO orbital energy (eV)
  N       number of atomic orbitals
  std     standard deviation of the last five SCF iteration energies (eV),
          the reference's convergence filter keeps std < 0.01 eV
  y       HOMO-LUMO gap (eV)
  pos     atom positions (Bohr), flattened [x0, y0, z0, x1, ...]

Its generation loop (pyscf_ipu/dft.py:621-657) writes one row per conformer
after the SCF, with energies converted by `hartree_to_eV` and the spread of the
last five iteration energies.  Here the rows come from the batched form
(`rows_from_batch`: thousands of molecules per launch) or the plain form
(`rows_from_plain`), and `write_parquet` / `read_parquet` store them with
pyarrow.  The library generates molecules from atom strings, not SMILES: a row
takes the caller's SMILES when given, else the Hill-order formula.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

from .gto import ELEMENTS
from .nanoDFT.nanoDFT import HARTREE_TO_EV

COLUMNS = ("smile", "atoms", "z", "energy", "homo", "lumo", "N", "std", "y", "pos")
CONVERGED_STD_EV = 0.01  # the reference's filter on `std` (qm1b/datasheet.pdf, section 1.3.2)

_SYMBOL = {z: s for s, z in ELEMENTS.items()}


def formula(z: Sequence[int]) -> str:
    """Hill-order formula (C first, H second, then alphabetical), e.g. "C2H6O"."""
    counts: Dict[str, int] = {}
    for a in z:
        s = _SYMBOL[int(a)]
        counts[s] = counts.get(s, 0) + 1
    order = [s for s in ("C", "H") if s in counts] + sorted(s for s in counts if s not in ("C", "H"))
    return "".join(s + (str(counts[s]) if counts[s] > 1 else "") for s in order)


def _row(z, pos_bohr, smile, energy_ha, homo_ha, lumo_ha, nao, last_energies_ha) -> dict:
    z = [int(a) for a in z]
    last = np.asarray([e for e in last_energies_ha if np.isfinite(e)], np.float64)
    return {"smile": smile if smile is not None else formula(z),
            "atoms": "".join(_SYMBOL[a] for a in z),
            "z": z,
            "energy": float(energy_ha) * HARTREE_TO_EV,
            "homo": float(homo_ha) * HARTREE_TO_EV,
            "lumo": float(lumo_ha) * HARTREE_TO_EV,
            "N": int(nao),
            "std": float(np.std(last * HARTREE_TO_EV)) if len(last) else float("nan"),
            "y": float(abs(lumo_ha - homo_ha)) * HARTREE_TO_EV,
            "pos": [float(x) for x in np.asarray(pos_bohr, np.float64).reshape(-1)]}


def rows_from_batch(b, result, smiles: Optional[Sequence[str]] = None, only_converged: bool = True) -> List[dict]:
    """Rows of a finished `BatchRKS` run (`result` = its BatchResult): HOMO and
    LUMO from each molecule's final orbital energies, `std` over the last five
    iteration energies the batched energy kernel kept.  Molecules that failed
    set-up, and (with only_converged) those that did not converge, are left out."""
    hist = b.energy_history()
    mo = b.mo_e.cpu().numpy()
    rows = []
    for m, (z, R) in enumerate(b.molecules):
        if result.failed[m] or (only_converged and not result.converged[m]):
            continue
        nocc, v0 = int(b.desc[m, 3]), int(b.desc[m, 13])
        rows.append(_row(z, R, smiles[m] if smiles is not None else None, result.energy[m], mo[v0 + nocc - 1],
                         mo[v0 + nocc], int(b.nao[m]), hist[m]))
    return rows


def rows_from_plain(mols, results, smiles: Optional[Sequence[str]] = None, only_converged: bool = True) -> List[dict]:
    """Rows of plain-form results (`nanoDFT.RKS.run` / `run_molecules`), one per
    molecule (`mols`: the Mole objects the results belong to)."""
    rows = []
    for m, (mol, r) in enumerate(zip(mols, results)):
        if only_converged and not r.converged:
            continue
        nocc = mol.nelectron // 2
        rows.append(_row(mol.atom_charges(), mol.atom_coords(), smiles[m] if smiles is not None else None,
                         r.energy, r.mo_energy[nocc - 1], r.mo_energy[nocc], mol.nao_nr(), r.energies[-5:, 0]))
    return rows


def to_table(rows: List[dict]):
    import pyarrow as pa
    schema = pa.schema([("smile", pa.string()), ("atoms", pa.string()), ("z", pa.list_(pa.int32())),
                        ("energy", pa.float64()), ("homo", pa.float64()), ("lumo", pa.float64()),
                        ("N", pa.int32()), ("std", pa.float64()), ("y", pa.float64()),
                        ("pos", pa.list_(pa.float64()))])
    cols = {c: [r[c] for r in rows] for c in COLUMNS}
    return pa.table(cols, schema=schema)


def write_parquet(rows: List[dict], path: str) -> None:
    import pyarrow.parquet as pq
    pq.write_table(to_table(rows), path)


def read_parquet(path: str):
    """The rows back as a pandas DataFrame (as the reference's README reads them)."""
    import pandas as pd
    return pd.read_parquet(path)
