"""Large molecules (6-31G, nao 130-200) and the precision study of the AO-grid
products.

`LargeRKS` is the RKS SCF of `nanoDFT.RKS` (same iteration, DIIS, eigh and
energy, pyscf_ipu/nanoDFT/nanoDFT.py:40-102, 338-400) with data paths sized
for large nao:

  * ERIs: packed s8 from `pscf_eri_s8` (FP32 or FP64), kept packed; J/K
    from them with rows of dm and of the accumulator in LDS
    (`pscf_jk_s8_large`, FP64 accumulation), 8x fewer bytes per iteration
    than streaming the N^4 s1 tensor (`pscf_jk_s1`, still in the library);
  * XC: the AO values (4, N, g) are stored in the study's input precision and
    the two AO-grid products X = D phi0 and V = phi0 Y^T run on `pscf_gemm`
    (BF16 / FP16 / FP32 / FP64 inputs, FP32 or FP64 accumulation); the
    pointwise parts (rho, B3LYP, Y) accumulate in FP64;
  * SCF step: `pscf_scf_step_ws` (working matrices in a global workspace).

    python -m pyscf_ipu_amd.precision --molecules 64 --out precision_sweep.json

runs the study: 64 generator molecules with 12-16 heavy atoms, 6-31G, grid
level 1, every (input, accumulation) variant against the FP64/FP64 run of the
same molecule with FP64 ERIs, reporting the errors of total energy, HOMO-LUMO
gap and density per molecule.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import Dict, List

import numpy as np
import torch

from . import _lib, device, grids, guess, molgen
from .electron_repulsion import direct
from .gto import build_mol
from .nanoDFT.nanoDFT import EPSILON_B3LYP, HYB_B3LYP, RKS, SCFResult

# (storage type of the AO values / density / Y, accumulate in FP64)
TYPES = {"bf16": (0, torch.bfloat16), "f16": (1, torch.float16), "f32": (2, torch.float32), "f64": (3, torch.float64)}
VARIANTS = [f"{t}/{a}" for t in ("bf16", "f16", "f32", "f64") for a in ("f32", "f64")]


class LargeRKS(RKS):
    """RKS for large nao.  `gemm` = "<input>/<accumulation>" of the AO-grid
    GEMMs, `eri` = "fp32" or "fp64" (ERI generation and s1 storage)."""

    NATIVE_LOOP = False  # its set-up, Fock phases and step are its own

    def __init__(self, mol, level: int = 1, gemm: str = "f64/f64", eri: str = "fp64", grid=None, events=None):
        super().__init__(mol, level=level, dtype="fp64", grid=grid, events=events)
        tin, acc = gemm.split("/")
        self.gemm_type, self.gemm_tdt = TYPES[tin]
        self.acc64 = int(acc == "f64")
        self.eri_dtype = eri
        self.gemm_name = gemm

    def setup(self, dm0=None):
        mol = self.mol
        t0 = time.perf_counter()
        device.require_gpu()
        f64 = torch.float64
        self.N = N = mol.nao_nr()
        self.nocc = mol.nelectron // 2
        atm, bas, env = direct._mol_device(mol)
        self.dev_mol = (atm, bas, env)
        S, T, V = device.empty((N, N), f64), device.empty((N, N), f64), device.empty((N, N), f64)
        _lib.call("pscf_int1e", device.ptr(atm), mol.natm, device.ptr(bas), mol.nbas, device.ptr(env), N,
                  device.ptr(S), device.ptr(T), device.ptr(V), device.stream())
        self.S = S
        self.Hcore = (T + V).contiguous()
        S_h = S.cpu().numpy()
        self.L_inv_h = np.linalg.inv(np.linalg.cholesky(S_h))
        self.L_inv = device.to_dev(self.L_inv_h)
        g = self.grid if self.grid is not None else grids.atomic_grids(mol, self.level)
        self.g = G = g.size
        self.coords = device.to_dev(g.coords)
        if g.weights is None:
            owner, vol = device.to_dev(g.owner.astype(np.int32)), device.to_dev(g.vol)
            axyz = device.to_dev(mol.atom_coords())
            adj = device.to_dev(grids.becke_adjust_table(mol.atom_charges()))
            w64 = device.empty(G, f64)
            _lib.call("pscf_becke_weights", device.ptr(self.coords), device.ptr(owner), device.ptr(vol), G,
                      device.ptr(axyz), device.ptr(adj), mol.natm, device.ptr(w64), device.stream())
        else:
            w64 = device.to_dev(g.weights)
        self.weights64 = w64
        ao64 = device.empty((4, N, G), f64)
        _lib.call("pscf_ao_eval", _lib.F64, device.ptr(atm), mol.natm, device.ptr(bas), mol.nbas, device.ptr(env), N,
                  device.ptr(self.coords), G, device.ptr(ao64), device.stream())
        self.phi = self._cast(ao64)
        del ao64
        # ERIs: packed s8 (J/K reads them packed)
        self.eri_code = _lib.F64 if self.eri_dtype == "fp64" else _lib.F32
        self.s8 = direct.eri_s8(mol, self.eri_dtype, dev_mol=self.dev_mol)
        self.jk_ws_b = _lib.lib().pscf_jk_s8_large_workspace_bytes(N)
        self.jk_ws = device.workspace(self.jk_ws_b)
        # buffers
        self.D = device.to_dev(guess.minao(mol, S_h) if dm0 is None else dm0, f64)
        self.Dg = device.empty((N, N), self.gemm_tdt)
        self.X = device.empty((N, G), f64)
        self.coef = device.empty((4, G), f64)
        self.Y = device.empty((N, G), self.gemm_tdt)
        self.nparts = _lib.lib().pscf_xc_large_parts(G)
        self.e_part = device.empty(self.nparts, f64)
        self.Vh = device.empty((N, N), f64)
        self.Vxc = device.empty((N, N), f64)
        self.dJK = device.empty((N, N), f64)
        self.exc = device.empty(1, f64)
        self.H, self.C = device.empty((N, N), f64), device.empty((N, N), f64)
        self.mo_e = device.empty(N, f64)
        self.E = device.empty(5, f64)
        self.state = device.zeros(_lib.lib().pscf_scf_state_bytes(N) // 8, f64)
        self.ws_b = _lib.lib().pscf_scf_step_ws_bytes(N)
        self.ws = device.workspace(self.ws_b)
        self.E_nuc = mol.energy_nuc()
        self._setup_done = True
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def _cast(self, x: torch.Tensor) -> torch.Tensor:
        out = device.empty(tuple(x.shape), self.gemm_tdt)
        _lib.call("pscf_cast_to", self.gemm_type, device.ptr(x), device.ptr(out), x.numel(), device.stream())
        return out

    def fock_parts(self):
        N, G, P, st = self.N, self.g, device.ptr, device.stream()
        _lib.call("pscf_cast_to", self.gemm_type, P(self.D), P(self.Dg), N * N, st)
        # X[nu][p] = sum_mu D[nu][mu] phi0[mu][p]
        _lib.call("pscf_gemm", self.gemm_type, self.acc64, 0, 0, 0, N, G, N, P(self.Dg), N, P(self.phi), G,
                  P(self.X), G, 1, st)
        _lib.call("pscf_xc_large_stage", 0, self.gemm_type, N, G, P(self.X), P(self.phi), P(self.weights64),
                  EPSILON_B3LYP, P(self.coef), P(self.e_part), None, st)
        _lib.call("pscf_xc_large_stage", 1, self.gemm_type, N, G, None, P(self.phi), None, EPSILON_B3LYP,
                  P(self.coef), None, P(self.Y), st)
        # Vh[mu][nu] = sum_p phi0[mu][p] Y[nu][p], K = G split over workgroups
        self.Vh.zero_()
        ksplit = max(1, min(256, G // 2048))
        _lib.call("pscf_gemm", self.gemm_type, self.acc64, 0, 1, 1, N, N, G, P(self.phi), G, P(self.Y), G,
                  P(self.Vh), N, ksplit, st)
        _lib.call("pscf_xc_large_finish", N, P(self.Vh), P(self.e_part), self.nparts, P(self.Vxc), P(self.exc), st)
        _lib.call("pscf_jk_s8_large", self.eri_code, P(self.s8), N, P(self.D), HYB_B3LYP, P(self.dJK), P(self.jk_ws),
                  self.jk_ws_b, st)

    def energy_terms(self):
        _lib.call("pscf_energy", _lib.F64, self.N, device.ptr(self.D), device.ptr(self.Hcore), device.ptr(self.dJK),
                  device.ptr(self.exc), float(self.E_nuc), device.ptr(self.E), device.stream())

    def _step(self, it: int, diis: bool = True):
        P = device.ptr
        _lib.call("pscf_scf_step_ws", _lib.F64, it, self.N, self.nocc, P(self.Hcore), P(self.S), P(self.L_inv),
                  P(self.dJK), P(self.Vxc), P(self.D), P(self.H), P(self.mo_e), P(self.C), P(self.state), int(diis),
                  P(self.ws), self.ws_b, device.stream())


def study_molecules(n: int = 64, seed: int = 4) -> List[str]:
    """`n` generator molecules with 12-16 heavy atoms (uniform)."""
    rng = np.random.default_rng(seed)
    return [molgen.to_mol_str(molgen.random_molecule(rng, int(rng.integers(12, 17)))) for _ in range(n)]


def sweep(mol_strs: List[str], variants=VARIANTS, level: int = 1, its: int = 60, conv_tol: float = 1e-8,
          eri_variants=("fp64", "fp32"), log=print) -> Dict:
    """Per molecule: FP64 reference (FP64 ERIs, f64/f64 GEMMs), then every GEMM
    variant with FP64 ERIs and the f64/f64 GEMMs with FP32 ERIs, each started
    from the reference's converged density.  Errors: |dE| (Ha), |d gap| (Ha),
    max |dD|."""
    out = {"variants": {}, "molecules": []}
    for i, s in enumerate(mol_strs):
        mol = build_mol(s, "6-31g")
        grid = grids.atomic_grids(mol, level)
        t0 = time.perf_counter()
        ref = LargeRKS(mol, level, "f64/f64", "fp64", grid=grid).run(its=its, conv_tol=conv_tol)
        rec = {"nao": mol.nao_nr(), "natm": mol.natm, "ngrid": grid.size, "E_ref": ref.energy,
               "ref_iterations": ref.iterations, "ref_converged": bool(ref.converged), "errors": {}}
        runs = [(v, "fp64") for v in variants if v != "f64/f64"] + [("f64/f64", e) for e in eri_variants
                                                                   if e != "fp64"]
        for gv, ev in runs:
            r = LargeRKS(mol, level, gv, ev, grid=grid).run(its=its, conv_tol=conv_tol, dm0=ref.dm)
            key = f"{gv} eri-{ev}"
            rec["errors"][key] = {"dE": abs(r.energy - ref.energy), "dgap": abs(r.hlgaps[-1] - ref.hlgaps[-1]),
                                  "dD": float(np.abs(r.dm - ref.dm).max()), "iterations": r.iterations}
        out["molecules"].append(rec)
        torch.cuda.empty_cache()
        log(f"precision: molecule {i + 1}/{len(mol_strs)} nao {rec['nao']} grid {rec['ngrid']} E {ref.energy:.6f} "
            f"({time.perf_counter() - t0:.1f} s) " + " ".join(f"{k}:{v['dE']:.1e}" for k, v in rec["errors"].items()))
    # statistics over the molecules whose FP64 reference converged
    ok = [m for m in out["molecules"] if m["ref_converged"]]
    out["n_ref_converged"] = len(ok)
    keys = ok[0]["errors"].keys() if ok else []
    for k in keys:
        a = np.array([[m["errors"][k]["dE"], m["errors"][k]["dgap"], m["errors"][k]["dD"]] for m in ok])
        out["variants"][k] = {"max_dE": float(a[:, 0].max()), "mean_dE": float(a[:, 0].mean()),
                              "median_dE": float(np.median(a[:, 0])), "max_dgap": float(a[:, 1].max()),
                              "median_dD": float(np.median(a[:, 2])), "max_dD": float(a[:, 2].max()),
                              "within_1mHa": int((a[:, 0] <= 1e-3).sum()), "n": len(a)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--molecules", type=int, default=64)
    ap.add_argument("--level", type=int, default=1)
    ap.add_argument("--out", default="precision_sweep.json")
    a = ap.parse_args(argv)
    torch.cuda.set_device(0)
    res = sweep(study_molecules(a.molecules), level=a.level, log=lambda s: print(s, flush=True))
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    for k, v in res["variants"].items():
        print(f"{k:18s} max|dE| {v['max_dE']:.2e} Ha  mean {v['mean_dE']:.2e}  median {v['median_dE']:.2e}  "
              f"max|dgap| {v['max_dgap']:.2e}  median|dD| {v['median_dD']:.2e}  max|dD| {v['max_dD']:.2e}  "
              f"within 1 mHa {v['within_1mHa']}/{v['n']}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
