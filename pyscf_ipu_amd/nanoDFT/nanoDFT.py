"""nanoDFT on the MI355X: restricted Kohn-Sham B3LYP SCF (drop-in for pyscf_ipu.nanoDFT.nanoDFT).

Kept names and meanings (pyscf_ipu/nanoDFT/nanoDFT.py):
  energy(density_matrix, H_core, diff_JK, E_xc, E_nuc)          :20-38
  exchange_correlation(density_matrix, grid_AO, grid_weights)   :105-123
  get_JK(density_matrix, ERI, dense_ERI, backend)               :125-139
  nanoDFT(mol, opts) -> (energies (its, 5) eV, (logged_energies, hlgaps,
                          mo_energy, mo_coeff, grid_coords, grid_weights))  :259-336
  hlgap(L_inv, H, n_electrons_half, _np)                        :366-368
  grad_nuc(charges, coords), grad(mol, coords, weight, mo_coeff,
           mo_energy) -> (natm, 3) Ha/Bohr                       :435-477
  structure_optimization (the loop main runs with the option)   :645-691
  nanoDFT_options(...)                                          :548-618
  main()                                                        :620-643

Per molecule everything after the host set-up runs in the library's kernels:
int1e, the Cholesky of S and L^-1, the atomic-density guess, ERIs (packed s8),
AO values, Becke weights, and per iteration the fused XC kernel, the s8 J/K
kernel, the one-workgroup SCF step (DIIS + Jacobi eigh + density) and the FP64
energy reduction.  The host keeps the grid templates and reads one energy per
iteration to stop at |dE| < conv_tol (the reference always runs `its`
iterations).
"""
from __future__ import annotations

import argparse
import ctypes
import time
from collections import namedtuple
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from .. import _lib, device, grids, guess
from ..electron_repulsion import direct
from ..gto import Mole, build_mol
from . import sparse_symmetric_ERI as ssE

HARTREE_TO_EV = 27.2114079527
EPSILON_B3LYP = 1e-20
HYB_B3LYP = 0.2


def energy(density_matrix, H_core, diff_JK, E_xc, E_nuc, _np=np):
    E_core = _np.sum(density_matrix * H_core)
    E_J_K = _np.sum(density_matrix * diff_JK)
    E = E_core + E_J_K / 2 + E_xc + E_nuc
    return _np.array([E, E_core, E_J_K / 2, E_xc, E_nuc])


def hlgap(L_inv, H, n_electrons_half, _np=np):
    mo_energy = _np.linalg.eigh(L_inv @ H @ L_inv.T)[0]
    return _np.abs(mo_energy[n_electrons_half] - mo_energy[n_electrons_half - 1])


def _dt(x):
    return torch.float64 if x.dtype == torch.float64 else torch.float32


def exchange_correlation(density_matrix, grid_AO, grid_weights, eps: float = EPSILON_B3LYP):
    """(E_xc, V_xc) for grid_AO in the reference's (4, g, N) layout (device or host)."""
    is_np = not isinstance(density_matrix, torch.Tensor)
    ao = grid_AO if isinstance(grid_AO, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(grid_AO))
    dt = _dt(ao)
    code = _lib.F64 if dt == torch.float64 else _lib.F32
    ao = device.to_dev(ao, dt)
    _, g, N = ao.shape
    dm = device.to_dev(density_matrix if not is_np else np.asarray(density_matrix), dt).reshape(N, N)
    w = device.to_dev(grid_weights if isinstance(grid_weights, torch.Tensor) else np.asarray(grid_weights), dt)
    exc = device.empty(1, torch.float64)
    vxc = device.empty((N, N), dt)
    wsb = _lib.lib().pscf_xc_workspace_bytes(code, N, g)
    ws = device.workspace(wsb)
    _lib.call("pscf_xc", code, device.ptr(ao), g * N, 1, N, g, N, device.ptr(dm), device.ptr(w), float(eps),
              device.ptr(exc), device.ptr(vxc), device.ptr(ws), wsb, device.stream())
    if is_np:
        return float(exc.item()), vxc.cpu().numpy()
    return exc[0], vxc


def get_JK(density_matrix, ERI, dense_ERI=False, backend="mi355x"):
    """dJK = J - K/2 * HYB_B3LYP (nanoDFT.py:125-139).  dense_ERI: ERI is the
    full (N, N, N, N) tensor (optionally with a leading axis of 1, as nanoDFT
    stacks it) and J = einsum('ijkl,ji->kl'), K = einsum('ijkl,jk->il') run in
    one streaming pass over it (`pscf_jk_s1`, FP64 accumulation; uses the 8-fold
    symmetry of real ERIs, any dm).  Otherwise ERI = (nonzero_distinct_ERI,
    nonzero_indices) as nanoDFT prepares them (`sparse_symmetric_einsum`).
    Returns (N, N) in the density's type: a device tensor, or numpy for numpy."""
    N = density_matrix.shape[0]
    if dense_ERI:
        is_np = not isinstance(density_matrix, torch.Tensor)
        eri = ERI if isinstance(ERI, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(ERI))
        if eri.numel() != N ** 4:
            raise ValueError(f"dense ERI has {eri.numel()} elements, expected N^4 = {N ** 4}")
        dt = _dt(eri)
        code = _lib.F64 if dt == torch.float64 else _lib.F32
        eri = device.to_dev(eri.reshape(-1), dt)
        dm = device.to_dev(density_matrix if not is_np else np.asarray(density_matrix), torch.float64).reshape(N, N)
        out = device.empty((N, N), torch.float64)
        _lib.call("pscf_jk_s1", code, device.ptr(eri), N, device.ptr(dm), HYB_B3LYP, device.ptr(out),
                  device.stream())
        odt = np.asarray(density_matrix).dtype if is_np else density_matrix.dtype
        if is_np:
            return out.cpu().numpy().astype(odt)
        return out.to(odt)
    out = ssE.sparse_symmetric_einsum(ERI[0], ERI[1], density_matrix, backend)
    return out.reshape(N, N)


# ------------------------------------------------------------------ driver ---
class KernelEvents:
    """HIP-event timing of named launch groups on the current stream.

    Records a start/stop event pair around each group; durations are read
    after the caller synchronises (`summary()`), so timing adds no host sync."""

    def __init__(self):
        self.pairs = {}
        self.meta = {}

    def record(self, name):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        self.pairs.setdefault(name, []).append([a, b])
        return b

    def summary(self):
        out = {}
        for k, v in self.pairs.items():
            ms = [a.elapsed_time(b) for a, b in v]
            out[k] = {"calls": len(ms), "total_ms": float(np.sum(ms)), "mean_ms": float(np.mean(ms))}
        return out


class _timed:
    def __init__(self, ev: Optional[KernelEvents], name: str):
        self.ev, self.name, self.stop = ev, name, None

    def __enter__(self):
        if self.ev is not None:
            self.stop = self.ev.record(self.name)
        return self

    def __exit__(self, *exc):
        if self.stop is not None:
            self.stop.record()
        return False


@dataclass
class SCFResult:
    energy: float                 # final total energy (Ha)
    energies: np.ndarray          # (n_iter, 5) [E, E_core, E_JK, E_xc, E_nuc] Ha
    hlgaps: np.ndarray            # (n_iter,) Ha
    mo_energy: np.ndarray
    mo_coeff: np.ndarray
    dm: np.ndarray
    iterations: int
    converged: bool
    timings: dict = field(default_factory=dict)


class RKS:
    """Plain form: one molecule, one sequence of launches per SCF phase.

    dtype "fp64": every kernel in FP64.  dtype "fp32": ERIs, J/K and XC in FP32;
    SCF step (DIIS, eigh, density) and the energy reduction in FP64."""

    # the library runs the whole loop of this class's own launches (subclasses
    # that replace a phase set this False)
    NATIVE_LOOP = True

    def __init__(self, mol: Mole, level: int = 0, dtype: str = "fp64", grid: Optional[grids.Grid] = None,
                 events: Optional["KernelEvents"] = None):
        self.mol = mol
        self.level = level
        self.dtype = dtype
        self.tdt, self.code, _ = device.DTYPES[dtype]
        self.grid = grid
        self.events = events
        self._setup_done = False

    # -- set-up ----------------------------------------------------------------
    def setup(self, dm0: Optional[np.ndarray] = None):
        mol = self.mol
        t0 = time.perf_counter()
        dev = device.require_gpu()
        self.N = N = mol.nao_nr()
        self.nocc = mol.nelectron // 2
        f64 = torch.float64
        # every host input in one upload (molecule, descriptor, grid, shell
        # pair lists): each separate copy from pageable memory costs ~0.1 ms
        g = self.grid if self.grid is not None else grids.atomic_grids(mol, self.level)
        self.g = g.size
        pairs, counts = direct.shell_pairs_by_class(mol)
        desc = np.zeros((1, 16), np.int64)
        desc[0, :4] = (mol.natm, mol.nbas, N, self.nocc)
        host = [mol._atm.astype(np.int32), mol._bas.astype(np.int32), mol._env.astype(np.float64), desc,
                np.asarray(g.coords, np.float64), pairs, mol.ao_loc_nr()[:-1].astype(np.int32)]
        if g.weights is None:
            host += [g.owner.astype(np.int32), np.asarray(g.vol, np.float64),
                     np.asarray(mol.atom_coords(), np.float64), grids.becke_adjust_table(mol.atom_charges())]
        else:
            host += [np.asarray(g.weights, np.float64)]
        up = device.upload_many(*host)
        atm, bas, env, self.d_desc, self.coords, d_pairs, shell_ao = up[:7]
        self.dev_mol = (atm, bas, env)
        S, T, V = device.empty((N, N), f64), device.empty((N, N), f64), device.empty((N, N), f64)
        _lib.call("pscf_int1e", device.ptr(atm), mol.natm, device.ptr(bas), mol.nbas, device.ptr(env), N,
                  device.ptr(S), device.ptr(T), device.ptr(V), device.stream())
        self.S = S
        self.Hcore = (T + V).contiguous()
        # overlap Cholesky -> L^-1 and the atomic-density guess on the device
        # (the batched kernels over a one-molecule descriptor)
        ctl = device.zeros((1, 8), f64)
        self.L_inv = device.empty((N, N), f64)
        _lib.call("pscf_batch_chol", device.ptr(self.d_desc), 1, N, device.ptr(S), device.ptr(self.L_inv),
                  device.ptr(ctl), device.stream())
        self._ctl1 = ctl
        # grid: atom-centred templates on the host, Becke weights on the device
        if g.weights is None:
            owner, vol, axyz, adj = up[7:11]
            w64 = device.empty(g.size, f64)
            _lib.call("pscf_becke_weights", device.ptr(self.coords), device.ptr(owner), device.ptr(vol), g.size,
                      device.ptr(axyz), device.ptr(adj), mol.natm, device.ptr(w64), device.stream())
        else:
            w64 = up[7]
        self.weights64 = w64
        self.weights = w64.to(self.tdt)
        self.ao = device.empty((4, N, g.size), self.tdt)
        _lib.call("pscf_ao_eval", self.code, device.ptr(atm), mol.natm, device.ptr(bas), mol.nbas, device.ptr(env), N,
                  device.ptr(self.coords), g.size, device.ptr(self.ao), device.stream())
        # ERIs, packed s8
        with _timed(self.events, "eri_s8"):
            self.eri = direct.eri_s8(mol, self.dtype, dev_mol=self.dev_mol, dev_pairs=(d_pairs, counts))
        # work buffers
        if dm0 is None:
            self.D = device.empty((N, N), f64)
            _lib.call("pscf_batch_guess", _lib.F64, device.ptr(self.d_desc), 1, device.ptr(atm), device.ptr(bas),
                      device.ptr(shell_ao), device.ptr(S), device.ptr(self.D), None, device.stream())
            self._shell_ao = shell_ao
        else:
            self.D = device.to_dev(dm0, f64)
        self.Dt = self.D.to(self.tdt) if self.tdt != f64 else self.D
        self.dJK = device.empty((N, N), self.tdt)
        self.Vxc = device.empty((N, N), self.tdt)
        self.exc = device.empty(1, f64)
        self.H = device.empty((N, N), f64)
        self.C = device.empty((N, N), f64)
        self.mo_e = device.empty(N, f64)
        self.E = device.empty(5, f64)
        self.state = device.zeros(_lib.lib().pscf_scf_state_bytes(N) // 8, f64)
        self.ws_xc_b = _lib.lib().pscf_xc_workspace_bytes(self.code, N, self.g)
        self.ws_xc = device.workspace(self.ws_xc_b)
        self.ws_jk_b = _lib.lib().pscf_jk_s8_workspace_bytes(self.code, N)
        self.ws_jk = device.workspace(self.ws_jk_b)
        self.E_nuc = mol.energy_nuc()
        self._setup_done = True
        # this stream only: molecules set up on other streams run on meanwhile
        torch.cuda.current_stream().synchronize()
        return time.perf_counter() - t0

    # -- kernels of one iteration ---------------------------------------------------
    def fock_parts(self):
        """V_xc, E_xc and dJK of the current density (self.D)."""
        N = self.N
        if self.tdt != torch.float64:
            _lib.call("pscf_cast", _lib.F64, device.ptr(self.D), self.code, device.ptr(self.Dt), N * N,
                      device.stream())
        with _timed(self.events, "xc"):
            _lib.call("pscf_xc", self.code, device.ptr(self.ao), N * self.g, self.g, 1, self.g, N,
                      device.ptr(self.Dt), device.ptr(self.weights), EPSILON_B3LYP, device.ptr(self.exc),
                      device.ptr(self.Vxc), device.ptr(self.ws_xc), self.ws_xc_b, device.stream())
        with _timed(self.events, "jk_s8"):
            _lib.call("pscf_jk_s8", self.code, device.ptr(self.eri), N, device.ptr(self.Dt), HYB_B3LYP,
                      device.ptr(self.dJK), device.ptr(self.ws_jk), self.ws_jk_b, device.stream())

    def energy_terms(self):
        _lib.call("pscf_energy", self.code, self.N, device.ptr(self.D), device.ptr(self.Hcore), device.ptr(self.dJK),
                  device.ptr(self.exc), float(self.E_nuc), device.ptr(self.E), device.stream())

    def step(self, it: int, diis: bool = True):
        with _timed(self.events, "scf_step"):
            self._step(it, diis)

    def _step(self, it: int, diis: bool = True):
        _lib.call("pscf_scf_step", self.code, it, self.N, self.nocc, device.ptr(self.Hcore), device.ptr(self.S),
                  device.ptr(self.L_inv), device.ptr(self.dJK), device.ptr(self.Vxc), device.ptr(self.D),
                  device.ptr(self.H), device.ptr(self.mo_e), device.ptr(self.C), device.ptr(self.state), int(diis),
                  device.stream())

    def _run_native(self, its, conv_tol, diis, t0, t1) -> SCFResult:
        N = self.N
        h_buf = _pinned(5 + N)
        log = np.zeros((max(its, 1), 6))
        n_it, conv = ctypes.c_int(0), ctypes.c_int(0)
        P = device.ptr
        _lib.call("pscf_rks_run", self.code, N, self.nocc, self.g, its, float(conv_tol), int(diis), P(self.ao),
                  P(self.weights), P(self.eri), P(self.Hcore), P(self.S), P(self.L_inv), P(self.D),
                  P(self.Dt) if self.tdt != torch.float64 else None, P(self.dJK), P(self.Vxc), P(self.exc),
                  P(self.H), P(self.mo_e), P(self.C), P(self.state), P(self.E), P(self.ws_xc), self.ws_xc_b,
                  P(self.ws_jk), self.ws_jk_b, float(self.E_nuc), ctypes.c_void_p(h_buf.data_ptr()),
                  log.ctypes.data_as(ctypes.c_void_p), ctypes.addressof(n_it), ctypes.addressof(conv), device.stream())
        n = n_it.value
        torch.cuda.current_stream().synchronize()
        t2 = time.perf_counter()
        return SCFResult(float(log[n - 1, 0]) if n else float(h_buf[0]), log[:n, :5].copy(), log[:n, 5].copy(),
                         self.mo_e.cpu().numpy(), self.C.cpu().numpy(), self.D.cpu().numpy(), n, bool(conv.value),
                         {"setup_s": t1 - t0, "scf_s": t2 - t1})

    def run(self, its: int = 20, conv_tol: float = 1e-8, diis: bool = True, dm0=None) -> SCFResult:
        """SCF to |E_i - E_(i-1)| < conv_tol or `its` iterations.  The loop runs
        in the library (pscf_rks_run) unless per-kernel events are requested,
        in which case it is issued from here, one launch group at a time."""
        t0 = time.perf_counter()
        if not self._setup_done:
            self.setup(dm0)
        t1 = time.perf_counter()
        if self.events is None and self.NATIVE_LOOP:
            return self._run_native(its, conv_tol, diis, t0, t1)
        self.fock_parts()
        self.energy_terms()
        E_prev = float(self.E[0].item())
        log, gaps = [], []
        converged = False
        for i in range(its):
            self.step(i, diis)
            self.fock_parts()
            self.energy_terms()
            e = self.E.cpu().numpy().copy()
            mo = self.mo_e.cpu().numpy()
            log.append(e)
            gaps.append(abs(mo[self.nocc] - mo[self.nocc - 1]))
            if abs(e[0] - E_prev) < conv_tol:
                converged = True
                break
            E_prev = e[0]
        torch.cuda.current_stream().synchronize()
        t2 = time.perf_counter()
        log = np.array(log)
        return SCFResult(float(log[-1, 0]), log, np.array(gaps), self.mo_e.cpu().numpy(), self.C.cpu().numpy(),
                         self.D.cpu().numpy(), len(log), converged, {"setup_s": t1 - t0, "scf_s": t2 - t1})


class DenseRKS(RKS):
    """RKS with the dense s1 ERI tensor (nanoDFT's `--dense_ERI`,
    nanoDFT.py:275-281, get_JK :128-131): the packed s8 ERIs are expanded once
    to N^4 values (`pscf_eri_s8_to_s1`) and dJK is one streaming pass over them
    per iteration (`pscf_jk_s1`, FP64 accumulation).  Everything else is RKS."""

    NATIVE_LOOP = False

    def setup(self, dm0=None):
        t = super().setup(dm0)
        N = self.N
        self.s1 = device.empty(N ** 4, self.tdt)
        _lib.call("pscf_eri_s8_to_s1", self.code, device.ptr(self.eri), self.code, device.ptr(self.s1), N,
                  device.stream())
        self.dJK64 = self.dJK if self.tdt == torch.float64 else device.empty((N, N), torch.float64)
        torch.cuda.current_stream().synchronize()
        return t

    def fock_parts(self):
        N = self.N
        if self.tdt != torch.float64:
            _lib.call("pscf_cast", _lib.F64, device.ptr(self.D), self.code, device.ptr(self.Dt), N * N,
                      device.stream())
        with _timed(self.events, "xc"):
            _lib.call("pscf_xc", self.code, device.ptr(self.ao), N * self.g, self.g, 1, self.g, N,
                      device.ptr(self.Dt), device.ptr(self.weights), EPSILON_B3LYP, device.ptr(self.exc),
                      device.ptr(self.Vxc), device.ptr(self.ws_xc), self.ws_xc_b, device.stream())
        with _timed(self.events, "jk_s1"):
            _lib.call("pscf_jk_s1", self.code, device.ptr(self.s1), N, device.ptr(self.D), HYB_B3LYP,
                      device.ptr(self.dJK64), device.stream())
            if self.dJK64 is not self.dJK:
                _lib.call("pscf_cast", _lib.F64, device.ptr(self.dJK64), self.code, device.ptr(self.dJK), N * N,
                          device.stream())


def _pinned(n: int) -> torch.Tensor:
    return torch.empty(n, dtype=torch.float64, pin_memory=True)


def run_molecules(mols, streams: int = 4, level: int = 0, dtype: str = "fp64", its: int = 20,
                  conv_tol: float = 1e-8, diis: bool = True, events: Optional[KernelEvents] = None):
    """SCF of several independent molecules, each with its own launches (one
    molecule per launch, as nanoDFT runs them), up to `streams` molecules in
    flight on their own HIP streams from as many host threads.  A molecule's
    SCF is latency-bound (its eigensolver is one workgroup), so concurrent
    molecules fill otherwise idle CUs.  Streams beyond the process's hardware
    queues (GPU_MAX_HW_QUEUES, HIP's default 4, read at HIP start-up) share
    queues and run one after another.
    Results come back in input order and equal the one-at-a-time results."""
    import threading
    out = [None] * len(mols)
    if streams <= 1:
        for i, m in enumerate(mols):
            out[i] = RKS(m, level=level, dtype=dtype, events=events).run(its=its, conv_tol=conv_tol, diis=diis)
        return out
    nxt, lock, errors = [0], threading.Lock(), []

    def worker():
        s = torch.cuda.Stream()
        try:
            with torch.cuda.stream(s):
                while True:
                    with lock:
                        i = nxt[0]
                        nxt[0] += 1
                    if i >= len(mols):
                        break
                    out[i] = RKS(mols[i], level=level, dtype=dtype, events=events).run(its=its, conv_tol=conv_tol,
                                                                                        diis=diis)
        except BaseException as e:  # re-raised on the calling thread
            errors.append(e)
        finally:
            s.synchronize()

    threads = [threading.Thread(target=worker) for _ in range(min(streams, len(mols)))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    return out


def nanoDFT(mol, opts):
    """Reference-shaped entry point (nanoDFT.py:259-336): returns (energies
    (n, 5) in eV, (logged_energies eV, hlgaps eV, mo_energy, mo_coeff,
    grid_coords, grid_weights)); n = opts.its unless opts.conv_tol stops
    earlier.  opts.dense_ERI selects the s1 ERI tensor (DenseRKS)."""
    dtype = "fp32" if getattr(opts, "float32", False) else "fp64"
    conv = getattr(opts, "conv_tol", 0.0)
    rks = (DenseRKS if getattr(opts, "dense_ERI", False) else RKS)(mol, level=opts.level, dtype=dtype)
    res = rks.run(its=opts.its, conv_tol=conv if conv else -1.0, diis=opts.diis)
    energies = res.energies * HARTREE_TO_EV
    grid_coords = rks.coords.cpu().numpy()
    grid_weights = rks.weights64.cpu().numpy()
    return energies, (energies.copy(), res.hlgaps * HARTREE_TO_EV, res.mo_energy, res.mo_coeff, grid_coords,
                      grid_weights)


# ------------------------------------------------------------------ gradients ---
def grad_nuc(charges, coords):
    """d E_nuc / d R, (natm, 3) Ha/Bohr (nanoDFT.py:435-449; host, natm^2 terms)."""
    charges = np.asarray(charges, dtype=np.float64)
    coords = np.asarray(coords, dtype=np.float64)
    natm = len(charges)
    diff = coords[None, :, :] - coords[:, None, :]
    r3 = np.linalg.norm(diff, axis=2) ** 3
    r3[np.arange(natm), np.arange(natm)] = np.inf
    return np.sum(-(charges[:, None] * charges[None, :])[:, :, None] * diff / r3[:, :, None], axis=0)


def grad(mol, coords, weight, mo_coeff, mo_energy):
    """Nuclear gradient of the RKS B3LYP energy, (natm, 3) Ha/Bohr, as
    nanoDFT.py:451-477 (no grid-response term: the grid `coords`, `weight` of
    the SCF stay fixed).  The reference builds int2e_ip1, int1e_ip* and the
    deriv-2 AO values with PySCF and contracts them with dm0 = 2 C_occ C_occ^T
    and dme0 = 2 C_occ e_occ C_occ^T (grad_elec :403-433).  Here dm0 and dme0
    are two FP64 GEMMs on the device and each integral class is contracted as
    it is generated: pscf_grad_int1e (overlap, kinetic, nuclear attraction and
    its operator derivative), pscf_grad_eri (derivative ERIs over the unique
    shell quartets with the J and exchange densities), pscf_grad_xc (AO second
    derivatives in the kernel, B3LYP potential); the nuclear repulsion term
    (grad_nuc) is added on the host."""
    device.require_gpu()
    N, natm = mol.nao_nr(), mol.natm
    nocc = mol.nelectron // 2
    mo_coeff = np.asarray(mo_coeff, dtype=np.float64)
    e_occ = np.asarray(mo_energy, dtype=np.float64)[:nocc]
    c_occ = device.to_dev(np.ascontiguousarray(mo_coeff[:, :nocc]))
    c2 = device.to_dev(np.ascontiguousarray(2.0 * mo_coeff[:, :nocc]))           # 2 C_occ
    c2e = device.to_dev(np.ascontiguousarray(2.0 * mo_coeff[:, :nocc] * e_occ))  # 2 C_occ e_occ
    f64 = torch.float64
    dm0, dme0 = device.empty((N, N), f64), device.empty((N, N), f64)
    P, st = device.ptr, device.stream()
    G_F64 = 3
    _lib.call("pscf_gemm", G_F64, 1, 0, 1, 0, N, N, nocc, P(c2), nocc, P(c_occ), nocc, P(dm0), N, 1, st)
    _lib.call("pscf_gemm", G_F64, 1, 0, 1, 0, N, N, nocc, P(c2e), nocc, P(c_occ), nocc, P(dme0), N, 1, st)
    atm, bas, env = direct._mol_device(mol)
    g_dev = device.zeros((natm, 3), f64)
    _lib.call("pscf_grad_int1e", P(atm), natm, P(bas), mol.nbas, P(env), N, P(dm0), P(dme0), P(g_dev), st)
    pairs, counts = direct.shell_pairs_by_class(mol)
    dp = device.to_dev(pairs)
    _lib.call("pscf_grad_eri", P(atm), natm, P(bas), mol.nbas, P(env), P(dp), (ctypes.c_int * 3)(*counts), N,
              P(dm0), HYB_B3LYP, P(g_dev), st)
    coords_d = device.to_dev(np.ascontiguousarray(coords, dtype=np.float64))
    w_d = device.to_dev(np.ascontiguousarray(weight, dtype=np.float64))
    g = len(w_d)
    wsb = _lib.lib().pscf_grad_xc_workspace_bytes(N, mol.nbas, g)
    ws = device.workspace(wsb)
    _lib.call("pscf_grad_xc", P(coords_d), P(w_d), g, P(atm), natm, P(bas), mol.nbas, P(env), N, P(dm0),
              EPSILON_B3LYP, P(ws), wsb, P(g_dev), st)
    out = g_dev.cpu().numpy()
    return out + grad_nuc(mol.atom_charges(), mol.atom_coords())


def structure_optimization(opts, steps: int = 20, mol_str: str = None):
    """The reference's structure-optimisation loop (nanoDFT.py:645-691, run
    when opts.structure_optimization is set): 6-31G, a slightly randomised
    water dimer (positions [[0,1,1], ..., [0,6,6]] Angstrom plus normal(0, 0.3)
    noise from numpy seed 42), `steps` times: nanoDFT, grad, then positions
    -= forces, with the reference's units (positions in Angstrom, gradient in
    Ha/Bohr).  The reference renders every step with mogli/imageio; here the
    energies (eV, last SCF iteration) and the positions of each step are
    returned.  `mol_str` replaces the starting geometry."""
    opts = opts._replace(basis="6-31G")
    if mol_str is None:
        p = np.array([[0, 1, 1], [0, 2, 2], [0, 3, 3], [0, 4, 4], [0, 5, 5], [0, 6, 6]], dtype=np.float64)
        np.random.seed(42)
        p = p + np.random.normal(0, 0.3, p.shape)
        A = ["H", "O", "H", "H", "O", "H"]
    else:
        parts = [q.split() for q in mol_str.split(";") if q.strip()]
        A = [q[0] for q in parts]
        p = np.array([[float(x) for x in q[1:4]] for q in parts])
    energies, positions = [], []
    for _ in range(steps):
        s = "; ".join(f"{a} {float(x[0])!r} {float(x[1])!r} {float(x[2])!r}" for a, x in zip(A, p))
        mol = build_mol(s, opts.basis)
        E, (_logged, _gaps, mo_energy, mo_coeff, grid_coords, grid_weights) = nanoDFT(mol, opts)
        energies.append(float(E[-1, 0]))
        positions.append(p.copy())
        forces = grad(mol, grid_coords, grid_weights, mo_coeff, mo_energy)
        p = p - forces
    return np.array(energies), np.array(positions)


_OPTION_FIELDS = ["its", "mol_str", "float32", "basis", "xc", "backend", "level", "multv", "intv", "threads",
                  "threads_int", "diis", "structure_optimization", "eri_threshold", "ao_threshold", "batches",
                  "ndevices", "dense_ERI", "v", "profile", "vis_num_error", "molecule_name", "conv_tol"]


def nanoDFT_options(its: int = 20, mol_str: str = "methane", float32: bool = False, basis: str = "sto-3g",
                    xc: str = "b3lyp", backend: str = "mi355x", level: int = 0, multv: int = 2, intv: int = 1,
                    threads: int = 1, threads_int: int = 1, diis: bool = True, structure_optimization: bool = False,
                    eri_threshold: float = 0.0, ao_threshold: float = 0.0, batches: int = 32, ndevices: int = 1,
                    dense_ERI: bool = False, v: bool = False, profile: bool = False, vis_num_error: bool = False,
                    molecule_name: str = None, conv_tol: float = 0.0):
    """Options namedtuple with the reference's field names (nanoDFT.py:548-571)
    plus `conv_tol` (0 = run all `its` iterations like the reference).  Returns
    (opts, mol_str); `mol_str` is an atom string ("C 0 0 0; H ...") or
    "methane"; molecules are never fetched from the network."""
    if molecule_name is None:
        molecule_name = mol_str
    if mol_str.strip().lower() == "methane":
        from ..gto import METHANE
        mol_str = METHANE
    if xc.lower() != "b3lyp":
        raise ValueError("only b3lyp is supported")
    if ndevices != 1:
        raise ValueError("one GPU per process: ndevices must be 1")
    vals = dict(its=its, mol_str=mol_str, float32=float32, basis=basis, xc=xc, backend=backend, level=level,
                multv=multv, intv=intv, threads=threads, threads_int=threads_int, diis=diis,
                structure_optimization=structure_optimization, eri_threshold=eri_threshold,
                ao_threshold=ao_threshold, batches=batches, ndevices=ndevices, dense_ERI=dense_ERI, v=v,
                profile=profile, vis_num_error=vis_num_error, molecule_name=molecule_name, conv_tol=conv_tol)
    opts = namedtuple("DFTOptionsImmutable", _OPTION_FIELDS)(**vals)
    return opts, mol_str


def main(argv=None):
    p = argparse.ArgumentParser(description="nanoDFT (MI355X): RKS B3LYP SCF")
    p.add_argument("--its", type=int, default=20)
    p.add_argument("--mol_str", default="methane")
    p.add_argument("--float32", action="store_true")
    p.add_argument("--basis", default="sto-3g")
    p.add_argument("--level", type=int, default=0)
    p.add_argument("--conv_tol", type=float, default=0.0)
    p.add_argument("--no-diis", dest="diis", action="store_false")
    p.add_argument("--structure_optimization", action="store_true")
    a = p.parse_args(argv)
    opts, mol_str = nanoDFT_options(its=a.its, mol_str=a.mol_str, float32=a.float32, basis=a.basis, level=a.level,
                                    diis=a.diis, conv_tol=a.conv_tol, structure_optimization=a.structure_optimization)
    if opts.structure_optimization:
        E, _pos = structure_optimization(opts)
        for i, e in enumerate(E):
            print(f"step {i:3d}  E = {e:.8f} eV")
        return E
    mol = build_mol(mol_str, opts.basis)
    print(f"Number of Atomic Orbitals\t{mol.nao_nr():15d}")
    print(f"Number of electrons\t{mol.nelectron:15d}")
    E, (logged, gaps, *_rest) = nanoDFT(mol, opts)
    for i, (e, g) in enumerate(zip(E[:, 0], gaps)):
        print(f"iter {i:3d}  E = {e / HARTREE_TO_EV:.10f} Ha   HL gap = {g:.6f} eV")
    return E


if __name__ == "__main__":
    main()
