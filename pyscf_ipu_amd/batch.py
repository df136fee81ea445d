"""Batched RKS B3LYP SCF: thousands of molecules of uneven size per launch.

Every SCF phase is one launch over all molecules (per angular class for the
ERIs), driven by per-molecule descriptors (csrc/batch.h) into concatenated
device buffers:

  set-up     grid expansion from element templates -> Becke weights -> AO values
             -> S and H_core -> Cholesky / L^-1 -> atomic-density guess -> s8 ERIs
  iteration  SCF step (DIIS + eigh + D) -> fused XC -> s8 J/K -> energy and
             per-molecule convergence control (converged molecules drop out
             of every later launch)

One iteration can be captured in a HIP graph and replayed.  The host
assembly is vectorised numpy over all atoms/shells (no per-molecule Python
loops beyond input parsing).  Precision: dtype "fp32" runs ERIs, J/K, XC and
the Becke weights in FP32 with FP64 SCF step, DIIS and energies; "fp64" runs
everything in FP64.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib, device, grids, gto

DIIS_N = 9
CTL = 8
ENERGY_HIST = 5  # iteration energies kept per molecule (pscf_batch_energy's ring)
SPLITS = (1, 2, 4, 8)   # workgroups per molecule tried by the autotuner
MAX_SPLIT = max(SPLITS)
MIN_WORKGROUPS = 512  # a small class of molecules is split further, to fill the 256 CUs


class PhaseTimer:
    """HIP events around named phases on the current stream; `ms()` reads them
    after the caller has synchronised (recording adds no host sync)."""

    def __init__(self):
        self.marks = []

    def __call__(self, name: str):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        self.marks.append((name, e))

    def ms(self) -> dict:
        out = {}
        for (name, a), (_, b) in zip(self.marks, self.marks[1:]):
            out[name] = out.get(name, 0.0) + a.elapsed_time(b)
        return out


def class_split(split: int, count: int) -> int:
    """Workgroups per molecule of an XC or J/K launch over `count` molecules."""
    return min(64, max(split, -(-MIN_WORKGROUPS // count)))
JK_THREADS = (256, 512)   # J/K workgroup sizes tried by the autotuner
# grid points per XC tile.  The autotuner varies only launch shapes that leave
# the result's bits unchanged (XC split, J/K workgroup size and split); a
# 64-point tile groups each point's density sums differently, and it also
# measured slower than 32 (29 vs 16 ms on the 4096-molecule set).
XC_TILES = {"fp32": (32,), "fp64": (32,)}
XC_CHUNK_PTS = 512  # points per chunk of pscf_batch_xc (PSCF_XC_CHUNK_PTS)
_TUNE_CACHE = {}        # tuning results per batch shape
CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def state_doubles(N: np.ndarray) -> np.ndarray:
    return 2 * DIIS_N * N * N + 100 + N * N + 2


def class_pairs(X: int, ns: np.ndarray, npp: np.ndarray) -> np.ndarray:
    return ns * (ns + 1) // 2 if X == 0 else (ns * npp if X == 1 else npp * (npp + 1) // 2)


def parse(mol_strs: Sequence[str]) -> List[Tuple[np.ndarray, np.ndarray]]:
    """Atom strings (Angstrom) -> [(Z, R Bohr)]."""
    out = []
    for s in mol_strs:
        atoms = gto.parse_atom_string(s)
        out.append((np.array([gto.ELEMENTS[a] for a, _ in atoms], dtype=np.int32),
                    np.array([x for _, x in atoms], dtype=np.float64) / gto.BOHR))
    return out


@dataclass
class BatchResult:
    energy: np.ndarray       # (M,) Ha
    terms: np.ndarray        # (M, 5)
    iterations: np.ndarray   # (M,)
    converged: np.ndarray    # (M,) bool
    failed: np.ndarray       # (M,) bool (overlap not positive definite)
    hlgap: np.ndarray        # (M,) Ha
    n_iter_launches: int


class BatchRKS:
    def __init__(self, molecules: Sequence[Tuple[np.ndarray, np.ndarray]], basis: str = "sto-3g", level: int = 0,
                 dtype: str = "fp32"):
        self.molecules = molecules
        self.basis = basis
        self.level = level
        self.dtype = dtype
        self.tdt, self.code, _ = device.DTYPES[dtype]
        self.split = {"xc": 1, "jk": 1}
        # FP32 XC evaluates the AO values in the kernel from the shells (no
        # stored 4 x nao x ngrid AO block); FP64 keeps the stored form, whose
        # tile pipeline at nao 65-80 leaves no LDS for the shell table
        self.ao_direct = dtype == "fp32"
        self.xc_tile = 32
        self.jk_threads = 256
        self.tuning = None
        self._assemble()

    # ------------------------------------------------------------ host assembly ---
    def _assemble(self):
        mols = self.molecules
        M = len(mols)
        Z = np.concatenate([z for z, _ in mols]).astype(np.int32)
        R = np.concatenate([r for _, r in mols]).astype(np.float64).reshape(-1, 3)
        natm = np.array([len(z) for z, _ in mols], dtype=np.int64)
        atom_mol = np.repeat(np.arange(M), natm).astype(np.int32)
        A = len(Z)
        # element templates: shells and grids
        elems = sorted(set(Z.tolist()))
        slot = {z: i for i, z in enumerate(elems)}
        env = [0.0] * gto.PTR_ENV_START
        tmpl_rows, tmpl_cnt = [], []
        for z in elems:
            rows = []
            for l, exps, coefs in gto.element_shells(gto.SYMBOLS[z], self.basis):
                pe = len(env)
                env.extend(exps.tolist())
                pc = len(env)
                env.extend(gto.normalized_coefficients(l, exps, coefs).tolist())
                rows.append([0, l, len(exps), 1, 0, pe, pc, 0])
            tmpl_rows.append(rows)
            tmpl_cnt.append(len(rows))
        maxsh = max(tmpl_cnt)
        tmpl = np.zeros((len(elems), maxsh, 8), dtype=np.int32)
        for i, rows in enumerate(tmpl_rows):
            tmpl[i, :len(rows)] = rows
        coord0 = len(env)
        env = np.concatenate([np.array(env), R.reshape(-1)])
        eslot = np.array([slot[z] for z in Z.tolist()], dtype=np.int32) if A else np.zeros(0, np.int32)
        atm = np.zeros((A, 6), dtype=np.int32)
        atm[:, 0] = Z
        atm[:, 1] = coord0 + 3 * np.arange(A)
        atm[:, 2] = 1
        nsh_atom = np.array(tmpl_cnt, dtype=np.int64)[eslot]
        sh_atom = np.repeat(np.arange(A), nsh_atom)
        sh_first = np.repeat(np.cumsum(nsh_atom) - nsh_atom, nsh_atom)
        sh_in_atom = np.arange(len(sh_atom)) - sh_first
        bas = tmpl[eslot[sh_atom], sh_in_atom].copy()
        bas[:, 0] = sh_atom
        l = bas[:, 1].astype(np.int64)
        width = 2 * l + 1
        sh_mol = atom_mol[sh_atom]
        nbas = np.bincount(sh_mol, minlength=M).astype(np.int64)
        nao = np.bincount(sh_mol, weights=width, minlength=M).astype(np.int64)
        bas0 = np.cumsum(nbas) - nbas
        atm0 = np.cumsum(natm) - natm
        ao_start = np.cumsum(nao) - nao
        shell_ao = (np.cumsum(width) - width - ao_start[sh_mol]).astype(np.int32)
        nelec = np.bincount(atom_mol, weights=Z, minlength=M).astype(np.int64)
        if np.any(nelec % 2):
            raise ValueError("closed-shell molecules only (even electron counts)")
        nocc = nelec // 2
        order_sh = np.lexsort((np.arange(len(l)), l, sh_mol))
        ns = np.bincount(sh_mol, weights=(l == 0), minlength=M).astype(np.int64)
        npp = np.bincount(sh_mol, weights=(l == 1), minlength=M).astype(np.int64)
        if np.any(l > 1):
            raise ValueError("only s and p shells are supported")
        P = nao * (nao + 1) // 2
        n_s8 = P * (P + 1) // 2
        eri0 = np.cumsum(n_s8) - n_s8
        mat0 = np.cumsum(nao * nao) - nao * nao
        vec0 = np.cumsum(nao) - nao
        st = state_doubles(nao)
        state0 = np.cumsum(st) - st
        # grid from element templates
        tm, toff, tn = [], [], []
        off = 0
        for z in elems:
            c, v = grids.atomic_grid(int(z), self.level)
            tm.append(np.hstack([c, v[:, None]]))
            toff.append(off)
            tn.append(len(v))
            off += len(v)
        self.tmpl_grid = np.vstack(tm)
        self.tmpl_off = np.array(toff, dtype=np.int32)
        self.tmpl_n = np.array(tn, dtype=np.int32)
        apts = self.tmpl_n[eslot].astype(np.int64)
        atom_pt0 = np.cumsum(apts) - apts
        ngrid = np.bincount(atom_mol, weights=apts, minlength=M).astype(np.int64)
        grid0 = np.cumsum(ngrid) - ngrid
        ao_sz = 4 * nao * ngrid
        ao0 = np.cumsum(ao_sz) - ao_sz
        desc = np.stack([natm, nbas, nao, nocc, atm0, bas0, ns, npp, bas0, ngrid, grid0, eri0, mat0, vec0, ao0, state0],
                        axis=1).astype(np.int64)
        self.M, self.A = M, A
        self.nao, self.ngrid, self.natm, self.nocc = nao, ngrid, natm, nocc
        self.atm, self.bas, self.env = atm, bas, env
        self.shell_ao, self.shl_sorted = shell_ao, order_sh.astype(np.int32)
        # first AO of each atom within its molecule, and its AO count
        self.atom_ao0 = shell_ao[np.cumsum(nsh_atom) - nsh_atom].astype(np.int32) if A else np.zeros(0, np.int32)
        self.atom_nao = np.bincount(sh_atom, weights=width, minlength=A).astype(np.int32)
        self.atom_elem, self.atom_pt0, self.atom_mol = eslot, atom_pt0, atom_mol
        self.desc = desc
        # shell-table bounds of the direct XC form: shells per molecule, primitives per shell
        self.nsh_max = int(nbas.max()) if M else 1
        self.kp = int(bas[:, 2].max()) if len(bas) else 1
        self.sizes = {"eri": int(n_s8.sum()), "mat": int((nao * nao).sum()), "vec": int(nao.sum()),
                      "ao": int(ao_sz.sum()), "state": int(st.sum()), "grid": int(ngrid.sum())}
        # work lists of the grid expansion (per atom) and of the six ERI classes
        # (per molecule): chunk counts here, the (owner, first) lists built on
        # the device at set-up (pscf_batch_chunks)
        self.grid_counts = apts
        self.eri_counts = []
        for X, Y in CLASSES:
            nx, ny = class_pairs(X, ns, npp), class_pairs(Y, ns, npp)
            self.eri_counts.append(nx * (nx + 1) // 2 if X == Y else nx * ny)
        self.order = np.argsort(-nao, kind="stable").astype(np.int32)
        self.max_nao = int(nao.max())
        # XC and J/K launches per class of nao rounded up to 16 (LDS sized per
        # class): (first position in `order`, count, largest nao of the class)
        npad = (nao[self.order] + 15) // 16
        self.nao_classes = []
        for v in np.unique(npad)[::-1]:
            pos = np.where(npad == v)[0]
            self.nao_classes.append((int(pos[0]), len(pos), int(nao[self.order[pos[0]]])))


    # ------------------------------------------------------------ device set-up ---
    @staticmethod
    def _work_list(counts: np.ndarray, size: int = 256) -> torch.Tensor:
        """(owner, first) int32 pairs covering range(counts[i]) in steps of
        `size`, built on the device from the owners' chunk counts."""
        n = (np.asarray(counts, np.int64) + size - 1) // size
        first = np.zeros(len(n) + 1, np.int64)
        np.cumsum(n, out=first[1:])
        total = int(first[-1])
        work = device.empty((max(total, 1), 2), torch.int32)
        d_first = device.to_dev(first)
        _lib.call("pscf_batch_chunks", device.ptr(d_first), len(n), size, device.ptr(work), total, device.stream())
        return work[:total]

    def setup(self, dm0: torch.Tensor = None, guess=None):
        """Device set-up.  The initial density is `dm0` (concatenated FP64
        matrices) if given, else what `guess(self)` writes into self.D and
        self.Dt (called once S, L^-1 and ctl exist), else the atomic guess."""
        dev = device.require_gpu()
        f64, T = torch.float64, self.tdt
        S = self.sizes
        self.d_desc = device.to_dev(self.desc)
        self.d_atm, self.d_bas, self.d_env = (device.to_dev(self.atm), device.to_dev(self.bas), device.to_dev(self.env))
        self.d_shell_ao, self.d_shl = device.to_dev(self.shell_ao), device.to_dev(self.shl_sorted)
        self.d_order = device.to_dev(self.order)
        self.d_enuc = device.empty(self.M, torch.float64)
        _lib.call("pscf_batch_enuc", device.ptr(self.d_desc), self.M, device.ptr(self.d_atm), device.ptr(self.d_env),
                  device.ptr(self.d_enuc), device.stream())
        # grid
        G = S["grid"]
        self.G = G
        self.coords = device.empty((G, 3), f64)
        self.vol = device.empty(G, f64)
        self.owner = device.empty(G, torch.int32)
        self.ptmol = device.empty(G, torch.int32)
        gw = self._work_list(self.grid_counts)
        ae, ap0, am = (device.to_dev(self.atom_elem), device.to_dev(self.atom_pt0), device.to_dev(self.atom_mol))
        toff, tn, tg = device.to_dev(self.tmpl_off), device.to_dev(self.tmpl_n), device.to_dev(self.tmpl_grid)
        P = device.ptr
        st = device.stream()
        tm = self.setup_timer = PhaseTimer()
        tm("grid")
        _lib.call("pscf_batch_grid", P(gw), gw.shape[0], P(ae), P(ap0), P(am), P(toff), P(tn), P(tg),
                  P(self.d_atm), P(self.d_env), P(self.coords), P(self.vol), P(self.owner), P(self.ptmol), st)
        self.w = device.empty(G, T)
        tm("becke")
        _lib.call("pscf_batch_becke", self.code, P(self.d_desc), P(self.ptmol), P(self.owner), P(self.coords),
                  P(self.vol), G, P(self.d_atm), P(self.d_env), P(self.w), None, st)
        if not self.ao_direct:
            self.ao = device.empty(S["ao"], T)
            tm("ao")
            _lib.call("pscf_batch_ao", self.code, P(self.d_desc), P(self.ptmol), P(self.coords), G, P(self.d_atm),
                      P(self.d_bas), P(self.d_env), P(self.ao), st)
        # one-electron part, orthogonaliser, guess
        nm = S["mat"]
        self.S, self.Hc, self.Linv = device.empty(nm, f64), device.empty(nm, f64), device.empty(nm, f64)
        self.ctl = device.zeros((self.M, CTL), f64)
        self.ctl[:, 1] = 1.0
        tm("int1e")
        _lib.call("pscf_batch_int1e", P(self.d_desc), self.M, P(self.d_atm), P(self.d_bas), P(self.d_env),
                  P(self.d_shell_ao), P(self.S), P(self.Hc), st)
        tm("chol_guess")
        _lib.call("pscf_batch_chol", P(self.d_desc), self.M, self.max_nao, P(self.S), P(self.Linv), P(self.ctl), st)
        self.D = device.empty(nm, f64)
        self.Dt = device.empty(nm, T)
        if guess is not None:
            guess(self)
        elif dm0 is None:
            _lib.call("pscf_batch_guess", self.code, P(self.d_desc), self.M, P(self.d_atm), P(self.d_bas),
                      P(self.d_shell_ao), P(self.S), P(self.D), P(self.Dt), st)
        else:
            self.D.copy_(dm0)
            self.Dt.copy_(dm0)
        # ERIs
        self.eri = device.empty(S["eri"], T)
        works = [self._work_list(c) for c in self.eri_counts]
        arr = (ctypes.c_void_p * 6)(*[w.data_ptr() for w in works])
        cnt = (ctypes.c_int * 6)(*[w.shape[0] for w in works])
        tm("eri_s8")
        # stored over their multiplicity as they are written, the form the J/K
        # contraction takes
        _lib.call("pscf_batch_eri_s8_div", self.code, P(self.d_desc), self.M, arr, cnt, P(self.d_atm),
                  P(self.d_bas), P(self.d_env), P(self.d_shell_ao), P(self.d_shl), P(self.eri), st)
        tm("end")
        # iteration buffers.  dJK is FP64; in FP32 mode it is built
        # incrementally, dJK += JK(D_new - D_old) with the change in FP32, so the
        # FP32 rounding shrinks with the density change and the energy can
        # converge below the FP32 noise of a full build.
        self.incremental = T != f64
        self.dJK, self.Vxc = device.zeros(nm, f64), device.empty(nm, T)
        self.dDt = device.empty(nm, T) if self.incremental else None
        # per-chunk XC buffers ([chunk][molecule of the class][max nao^2]),
        # sized for the class that needs the most
        ngrid = self.desc[:, 9]
        nchunk = [int(-(-ngrid[self.order[s0:s0 + c]].max() // XC_CHUNK_PTS)) for s0, c, _ in self.nao_classes]
        self.xc_parts_n = max(k * c * n * n for k, (_, c, n) in zip(nchunk, self.nao_classes))
        self.xc_parts_e_n = max(k * c for k, (_, c, _) in zip(nchunk, self.nao_classes))
        self.xc_parts = device.empty(self.xc_parts_n, T)
        self.xc_parts_e = device.empty(self.xc_parts_e_n, f64)
        self.exc = device.empty(self.M, f64)
        self.H, self.C = device.empty(nm, f64), device.empty(nm, f64)
        self.mo_e = device.empty(S["vec"], f64)
        self.E = device.empty((self.M, 5), f64)
        # the last ENERGY_HIST iteration energies of each molecule (ring, slot = iteration % ENERGY_HIST)
        self.E_hist = device.zeros((self.M, ENERGY_HIST), f64)
        self.state = device.zeros(S["state"], f64)
        self._keep = [gw, ae, ap0, am, toff, tn, tg] + works
        torch.cuda.synchronize()

    # ---------------------------------------------------------------- iterations ---
    def _xc(self, ctl_ptr, split, tile, vxc=None, exc=None):
        P, st = device.ptr, device.stream()
        for start, count, cmax in self.nao_classes:
            order = ctypes.c_void_p(self.d_order.data_ptr() + 4 * start)
            out = (P(self.Vxc if vxc is None else vxc), P(self.exc if exc is None else exc),
                   class_split(split, count), tile, P(self.xc_parts), P(self.xc_parts_e), self.xc_parts_n,
                   self.xc_parts_e_n, st)
            if self.ao_direct:
                _lib.call("pscf_batch_xc_direct", self.code, P(self.d_desc), order, count, cmax, ctl_ptr,
                          P(self.coords), P(self.d_atm), P(self.d_bas), P(self.d_env), self.nsh_max, self.kp,
                          P(self.Dt), P(self.w), 1e-20, *out)
            else:
                _lib.call("pscf_batch_xc", self.code, P(self.d_desc), order, count, cmax, ctl_ptr, P(self.ao),
                          P(self.Dt), P(self.w), 1e-20, *out)

    def _jk(self, ctl_ptr, split, threads, dm, out=None):
        P, st = device.ptr, device.stream()
        for start, count, cmax in self.nao_classes:
            order = ctypes.c_void_p(self.d_order.data_ptr() + 4 * start)
            _lib.call("pscf_batch_jk_s8", self.code, P(self.d_desc), order, count, cmax, ctl_ptr, P(self.eri), P(dm),
                      0.2, P(self.dJK if out is None else out), int(self.incremental), class_split(split, count),
                      threads, st)

    def _fock(self, ctl_ptr, first=False):
        self._xc(ctl_ptr, self.split["xc"], self.xc_tile)
        dm = self.Dt if (first or not self.incremental) else self.dDt
        self._jk(ctl_ptr, self.split["jk"], self.jk_threads, dm)

    def tune(self, repeats: int = 2):
        """Autotuner: times the XC launch for every (tile, split) candidate and
        the J/K launch for every (workgroup size, split) candidate on this
        batch (all molecules active, current density, scratch outputs), keeps
        the fastest, and caches the choice for batches of the same shape.
        `self.tuning` records the timings (ms per launch) next to the fixed
        tiling (32-point tiles, 256-thread J/K workgroups, one workgroup per
        molecule)."""
        key = (self.dtype, self.M, self.max_nao, self.sizes["eri"], self.sizes["grid"], self.ao_direct)
        if key in _TUNE_CACHE:
            self.tuning = dict(_TUNE_CACHE[key], cached=True)
            self.split = dict(self.tuning["chosen_split"])
            self.xc_tile = self.tuning["chosen_tile"]
            self.jk_threads = self.tuning["chosen_jk_threads"]
            return self.tuning
        P = device.ptr
        ctl = P(self.ctl)
        vxc, exc = torch.empty_like(self.Vxc), torch.empty_like(self.exc)
        out = torch.zeros_like(self.dJK)
        cands = {"xc": [(t, k) for t in XC_TILES[self.dtype] for k in SPLITS[:3]],
                 "jk": [(t, k) for t in JK_THREADS for k in SPLITS]}
        fns = {"xc": lambda c: self._xc(ctl, c[1], c[0], vxc, exc),
               "jk": lambda c: self._jk(ctl, c[1], c[0], self.Dt, out)}
        names = {"xc": lambda c: f"tile{c[0]}/split{c[1]}", "jk": lambda c: f"threads{c[0]}/split{c[1]}"}
        res = {"xc": {}, "jk": {}}
        for name in ("xc", "jk"):
            for c in cands[name]:
                fns[name](c)  # warm (sets kernel attributes)
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                for _ in range(repeats):
                    fns[name](c)
                b.record()
                b.synchronize()
                res[name][names[name](c)] = a.elapsed_time(b) / repeats
        best_xc = min(cands["xc"], key=lambda c: res["xc"][names["xc"](c)])
        best_jk = min(cands["jk"], key=lambda c: res["jk"][names["jk"](c)])
        self.xc_tile, self.jk_threads = best_xc[0], best_jk[0]
        self.split = {"xc": best_xc[1], "jk": best_jk[1]}
        self.tuning = {"ms": res, "chosen_split": dict(self.split), "chosen_tile": self.xc_tile,
                       "chosen_jk_threads": self.jk_threads,
                       "fixed": {"xc": "tile32/split1", "jk": "threads256/split1"}}
        _TUNE_CACHE[key] = dict(self.tuning)
        return self.tuning

    def _energy(self, init, tol, max_its):
        P, st = device.ptr, device.stream()
        _lib.call("pscf_batch_energy", self.code, P(self.d_desc), self.M, P(self.d_enuc), P(self.D), P(self.Hc),
                  P(self.dJK), P(self.exc), P(self.mo_e), P(self.ctl), P(self.E), P(self.E_hist), int(init), float(tol),
                  int(max_its), st)

    def _iteration(self, tol, max_its, diis=True):
        P, st = device.ptr, device.stream()
        _lib.call("pscf_batch_scf_step", self.code, P(self.d_desc), P(self.d_order), self.M, self.max_nao, P(self.ctl),
                  P(self.Hc), P(self.S), P(self.Linv), P(self.dJK), P(self.Vxc), P(self.D), P(self.Dt), P(self.dDt),
                  P(self.H), P(self.mo_e), P(self.C), P(self.state), int(diis), st)
        self._fock(P(self.ctl))
        self._energy(False, tol, max_its)

    def run(self, max_its: int = 20, conv_tol: float = 1e-8, graph: bool = True, check_every: int = 2,
            diis: bool = True, tune: bool = False) -> BatchResult:
        if not hasattr(self, "eri"):
            self.setup()
        if tune:
            self.tune()
        self._fock(device.ptr(self.ctl), first=True)
        self._energy(True, conv_tol, max_its)
        launches = 0
        g = None
        if graph:
            # first iteration eagerly (sets kernel attributes), then capture one iteration
            self._iteration(conv_tol, max_its, diis)
            launches += 1
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._iteration(conv_tol, max_its, diis)
        while launches < max_its:
            if g is not None:
                g.replay()
            else:
                self._iteration(conv_tol, max_its, diis)
            launches += 1
            if launches % check_every == 0 or launches == max_its:
                if float(self.ctl[:, 1].sum().item()) == 0.0:
                    break
        torch.cuda.synchronize()
        ctl = self.ctl.cpu().numpy()
        E = self.E.cpu().numpy()
        return BatchResult(E[:, 0].copy(), E, ctl[:, 6].astype(int), ctl[:, 2] > 0, ctl[:, 2] < 0, ctl[:, 5].copy(),
                           launches)

    def restart(self, molecules: np.ndarray):
        """Prepares molecules (indices) for another `run` from the atomic guess,
        in place: their density, dJK accumulator, DIIS / warm-start state and
        control words are reset, and every other molecule stays inactive.
        Grids, integrals, ERIs and the tuned tiling are kept.  A molecule whose
        set-up failed (ctl[m, 2] < 0: S not positive definite, or over-size)
        has no valid L^-1 and cannot be restarted: ValueError."""
        idx = np.unique(np.asarray(molecules, dtype=np.int64))
        if len(idx) == 0:
            return
        failed = idx[self.ctl[device.to_dev(idx), 2].cpu().numpy() < 0]
        if len(failed):
            raise ValueError(f"restart: molecules {failed.tolist()[:8]} failed set-up and cannot be restarted")
        P, st = device.ptr, device.stream()
        nm = self.sizes["mat"]
        D0, Dt0 = device.empty(nm, torch.float64), device.empty(nm, self.tdt)
        _lib.call("pscf_batch_guess", self.code, P(self.d_desc), self.M, P(self.d_atm), P(self.d_bas),
                  P(self.d_shell_ao), P(self.S), P(D0), P(Dt0), st)

        def spans(start, length):  # element indices of [start, start + length) per molecule
            length = length.astype(np.int64)
            return np.repeat(start - (np.cumsum(length) - length), length) + np.arange(length.sum())

        e = device.to_dev(spans(self.desc[idx, 12], self.nao[idx] ** 2))
        self.D[e] = D0[e]
        self.Dt[e] = Dt0[e]
        self.dJK[e] = 0.0
        self.E_hist[device.to_dev(idx)] = 0.0
        self.state[device.to_dev(spans(self.desc[idx, 15], state_doubles(self.nao[idx])))] = 0.0
        ctl = self.ctl.cpu().numpy()
        ctl[:, 1] = 0.0
        ctl[idx, :] = 0.0
        ctl[idx, 1] = 1.0
        self.ctl.copy_(torch.from_numpy(ctl))
        torch.cuda.synchronize()

    # ---------------------------------------------------------------- accessors ---
    def energy_history(self) -> np.ndarray:
        """(M, ENERGY_HIST) the last iteration energies of each molecule in
        iteration order (oldest first); NaN where a molecule ran fewer."""
        h = self.E_hist.cpu().numpy()
        its = self.ctl[:, 0].cpu().numpy().astype(np.int64)
        out = np.full_like(h, np.nan)
        for k in range(ENERGY_HIST):
            it = its - (ENERGY_HIST - 1 - k)  # iteration number held in column k
            ok = it >= 0
            out[ok, k] = h[np.where(ok)[0], it[ok] % ENERGY_HIST]
        return out

    def matrix(self, buf: torch.Tensor, m: int) -> np.ndarray:
        N, o = int(self.nao[m]), int(self.desc[m, 12])
        return buf[o:o + N * N].reshape(N, N).cpu().numpy()
