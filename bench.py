"""Throughput benchmark of the RKS B3LYP SCF hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--dtype fp64|fp32]

Workload (one "step"): the plain method, one molecule per launch sequence:
32 seeded synthetic neutral closed-shell molecules with 9 heavy atoms (C/N/O),
STO-3G (nao 54-65), B3LYP, grid level 0, SCF from the minao-style guess to
|dE| < 1e-8 Ha (at most 20 iterations).  The timed region covers every
molecule's set-up (int1e, grid, AO values, ERIs, guess) and SCF, with the
molecule descriptions already parsed.  A molecule counts towards the metric
only if its total energy is within 1 mHa of the FP64 CPU oracle's
(tests/golden/plain_sto3g_l0.json).

Prints one JSON line (rank 0) with the metric, a roofline record for the
dominant kernel and a CPU baseline (the FP64 oracle on a bounded sample, timed
on this host's cores before the GPU is used).  The per-kernel times behind the
roofline come from HIP events on the launch stream in a separate pass over the
same molecules, one molecule at a time (events around every launch group would
perturb the timed region); `serial_one_stream` is that pass's throughput.

Beside the plain headline, `forms` holds the batched FP32 forms (every SCF
phase one launch over all molecules, HIP-graph iterations, learned or atomic
guess) on two workloads: 4096 molecules with 9 heavy atoms from the same
generator as the headline (equal work per molecule: its `speedup_vs_plain` is
the ratio to the headline), and 4096 molecules with 1-9 heavy atoms.  Every
molecule is gated: |E - E_fp64| <= 1 mHa against an FP64 batched run of the
same molecules, and a molecule counts only if both its FP32 run and its FP64
reference converged.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "molecules/sec at |ΔE|≤1 mHa vs FP64; mean SCF iters; batched+learned vs plain"
GOLDEN = os.path.join(ROOT, "tests", "golden", "plain_sto3g_l0.json")
HBM_PEAK_GBS = 8000.0
# plain form: molecules in flight across worker processes, each running one
# molecule at a time with its own launches (8 processes measured 25-30 % above
# 8 host threads of one process on the same box: the threads serialise on the
# interpreter lock during each molecule's set-up)
PLAIN_WORKERS = 8
# when worker processes cannot be started (a profiler attached to this process):
# host threads of this process, each on its own HIP stream and hardware queue
PLAIN_STREAMS = 8
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < PLAIN_STREAMS:  # read when HIP starts
    os.environ["GPU_MAX_HW_QUEUES"] = str(PLAIN_STREAMS)
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}


def _oracle_one(mol_str: str) -> float:
    """One molecule through the FP64 oracle (C Rys ERIs, int1e, grid, numpy
    SCF): seconds of wall time."""
    from oracle import rks
    from pyscf_ipu_amd import grids
    from pyscf_ipu_amd.gto import build_mol
    t0 = time.perf_counter()
    mol = build_mol(mol_str)
    rks.scf(mol, grid=grids.build_grid(mol, 0), its=20, conv_tol=1e-8)
    return time.perf_counter() - t0


def _host_info(cores: int) -> dict:
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    blas = None
    try:
        cfg = np.show_config(mode="dicts")
        b = cfg.get("Build Dependencies", {}).get("blas", {})
        blas = f"{b.get('name')} {b.get('version', '')}".strip()
    except Exception:
        pass
    import scipy
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "affinity_cores": cores, "blas": blas,
            "numpy": np.__version__, "scipy": scipy.__version__,
            "pyscf": "not importable (not installed; never fetched)"}


def _cpu_baseline(mol_strs):
    """The FP64 oracle on this host's cores, before anything uses the GPU.
    Mode A (the value): a process pool with one molecule per process, BLAS and
    OpenMP threads = 1, one process per core, over the first `cores` molecules
    of the set (one round).  Mode B: one molecule with all cores (OpenMP ERIs,
    threaded BLAS), seconds per molecule."""
    import multiprocessing as mp
    cores = len(os.sched_getaffinity(0))
    nproc = min(cores, len(mol_strs))
    keep = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in keep:
        os.environ[k] = "1"
    try:
        ctx = mp.get_context("spawn")  # fresh interpreters: the thread settings apply from their start
        t0 = time.perf_counter()
        with ctx.Pool(nproc) as pool:
            per = pool.map(_oracle_one, mol_strs[:nproc], chunksize=1)
        wall_a = time.perf_counter() - t0
    finally:
        for k, v in keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    os.environ["OMP_NUM_THREADS"] = str(cores)
    with ctx.Pool(1) as pool:  # all-core run in a fresh process (OpenMP reads its thread count at start)
        lat_b = pool.map(_oracle_one, mol_strs[nproc:nproc + 1] or mol_strs[:1])[0]
    if keep["OMP_NUM_THREADS"] is None:
        os.environ.pop("OMP_NUM_THREADS", None)
    else:
        os.environ["OMP_NUM_THREADS"] = keep["OMP_NUM_THREADS"]
    # steady-state rate: every core has finished its molecule after max(per)
    # seconds (the pool's start-up, interpreter and imports, is left out)
    return {"value": nproc / max(per), "unit": "molecules/s", "cores": nproc, "kind": "port",
            "sample": f"{nproc} of the benchmark's 9-heavy-atom molecules, one per process on {nproc} cores "
                      f"(BLAS/OpenMP threads 1), FP64 oracle: C Rys ERIs + numpy SCF to |dE| < 1e-8; "
                      f"{min(per):.1f}-{max(per):.1f} s per molecule, {wall_a:.1f} s wall with pool start-up",
            "mode_b_latency_s": lat_b, "mode_b": f"one molecule with {cores} OpenMP/BLAS threads",
            "host": _host_info(cores)}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_bench.md")
N_CU = 256


def _profile_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0] == kernel:
            return int(cells[4].replace(",", ""))
    return None


def _roof(bound, achieved, peak, unit, ms, model):
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit, "frac": achieved / peak,
            "mean_launch_ms": ms, "work_model": model}


BATCH_STEPS, BATCH_WARMUP = 3, 1  # timed / untimed repeats of each batched form (each one a whole batch)


def _fp64_reference(mols, chunk: int = 1024):
    """FP64 batched SCF of the same molecules (pinned to the oracle by the GPU
    tests), in chunks of `chunk` molecules to bound device memory: energies,
    gaps, convergence flags and iteration counts."""
    import torch
    from pyscf_ipu_amd import batch
    E, gap, conv, its = [], [], [], []
    for k in range(0, len(mols), chunk):
        ref = batch.BatchRKS(mols[k:k + chunk], dtype="fp64")
        r = ref.run(max_its=60, conv_tol=1e-9, graph=True)
        E.append(r.energy)
        gap.append(r.hlgap)
        conv.append(r.converged)
        its.append(r.iterations)
        del ref
        torch.cuda.empty_cache()
    return np.concatenate(E), np.concatenate(gap), np.concatenate(conv), np.concatenate(its)


def _batched(mol_strs, workload: str, steps: int, warmup: int, its: int, golden=None):
    """Batched forms on `mol_strs`, every SCF phase one launch over all
    molecules, FP32 ERI/JK/XC with FP64 eigh/DIIS/energy, HIP-graph captured
    iterations:
      learned  SCF from the trained pair network's guess; molecules not
               converged after 15 iterations, or converged with a HOMO-LUMO gap
               below the re-check threshold, re-run from the atomic guess;
      minao    SCF from the atomic guess (what the network is compared with).
    Gate, per molecule: its FP32 run and its FP64 batched reference both
    converged and |E - E_fp64| <= 1 mHa.  `golden` (oracle energies of the
    first molecules) adds the same gate against the CPU oracle."""
    import torch
    from pyscf_ipu_amd import batch, learned_guess as lg
    mols = batch.parse(mol_strs)
    nmol = len(mols)
    E64, gap64, conv64, its64 = _fp64_reference(mols)
    model = lg.PairMLP.load() if os.path.exists(lg.DEFAULT_WEIGHTS) else None
    out = {}
    for name, mdl in (("learned", model), ("minao", None)):
        if name == "learned" and mdl is None:
            continue

        def once():
            return lg.solve(mols, mdl, dtype="fp32", max_its=its, learned_its=15, conv_tol=1e-6, graph=True, tune=True)

        for _ in range(min(warmup, BATCH_WARMUP)):
            once()
            torch.cuda.empty_cache()
        reps = max(1, min(steps, BATCH_STEPS))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            g = once()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        torch.cuda.empty_cache()
        res = g.result
        dE = np.abs(res.energy - E64)
        ok = (dE <= 1e-3) & res.converged & ~res.failed & conv64
        miss = np.where(~ok)[0]
        rec = {"value": float(ok.sum() / wall), "unit": "molecules/s", "n_molecules": nmol,
               "ms_per_step": wall * 1e3, "timed_repeats": reps, "passed_1mha": int(ok.sum()),
               "missed": int(nmol - ok.sum()),
               "missed_fp64_reference_not_converged": int((~conv64).sum()),
               "missed_fp32_not_converged": int((~res.converged & conv64).sum()),
               "missed_above_1mha": int(((dE > 1e-3) & res.converged & conv64).sum()),
               "max_abs_dE_vs_fp64_ha": float(dE[conv64].max()),
               "mean_abs_dE_vs_fp64_ha": float(dE[conv64].mean()),
               "max_abs_dgap_vs_fp64_ha": float(np.abs(res.hlgap - gap64)[ok].max()) if ok.any() else None,
               "mean_scf_iters": float(g.iterations_total.mean()),
               "fallback_to_minao": int(g.fallback.sum()),
               "fallback_not_converged": int((g.fallback & ~g.small_gap).sum()),
               "fallback_small_gap": int(g.small_gap.sum()),
               "misses": [{"molecule": int(i), "abs_dE_ha": float(dE[i]), "converged": bool(res.converged[i]),
                           "fp64_converged": bool(conv64[i]), "fallback": bool(g.fallback[i]),
                           "gap_ha": float(res.hlgap[i]), "gap_fp64_ha": float(gap64[i]),
                           "iterations": int(g.iterations_total[i])} for i in miss[:8]],
               "autotune_ms_per_launch": {k: {a: round(b, 3) for a, b in v.items()}
                                          for k, v in (g.tuning or {}).get("ms", {}).items()},
               "autotune_chosen": {"xc_tile": (g.tuning or {}).get("chosen_tile"),
                                   "jk_threads": (g.tuning or {}).get("chosen_jk_threads"),
                                   "split": (g.tuning or {}).get("chosen_split")}}
        if name == "learned":
            rec["learned_guess_ceiling"] = (
                "pair-only features: from the learned start the SCF takes more iterations than from MINAO "
                "(1024 held-out 9-heavy-atom molecules, no re-runs: 13.9 against 11.9), so this form "
                "cannot beat the MINAO form")
        if golden is not None:
            n = len(golden)
            dO = np.abs(res.energy[:n] - np.asarray(golden))
            rec["oracle_gate"] = {"molecules": n, "passed_1mha": int(np.sum(dO <= 1e-3)),
                                  "max_abs_dE_vs_oracle_ha": float(dO.max())}
        out[name] = rec
    # per-kernel rooflines of the batched launches over all molecules: J/K and
    # XC from the autotuner's per-launch times of the chosen tiling (every
    # molecule active), the ERIs from the set-up events
    from pyscf_ipu_amd import roofline
    sizes_b = batch.BatchRKS(mols, dtype="fp32")
    tun = g.tuning or {}
    ms = tun.get("ms", {})
    rl = {}
    if ms:
        jk_ms = ms["jk"][f"threads{tun['chosen_jk_threads']}/split{tun['chosen_split']['jk']}"]
        xc_ms = ms["xc"][f"tile{tun['chosen_tile']}/split{tun['chosen_split']['xc']}"]
        jk_b = 4.0 * sizes_b.sizes["eri"]
        nao_b, ng_b = sizes_b.desc[:, 2].astype(np.float64), sizes_b.desc[:, 9].astype(np.float64)
        xc_fl = float(np.sum(4.0 * nao_b * nao_b * ng_b))
        rl["jk_s8"] = _roof("hbm", jk_b / (jk_ms * 1e-3) / 1e9, HBM_PEAK_GBS, "GB/s", jk_ms,
                            "4 B per unique s8 ERI of all molecules")
        rl["xc"] = _roof("mfma", xc_fl / (xc_ms * 1e-3) / 1e12, PEAK_TFLOPS["fp32"], "TFLOP/s", xc_ms,
                         "FP32 MFMA work 4 N^2 flops per grid point (X = phi0 D and V += phi0^T Y); the AO "
                         "values are evaluated in the kernel, so no AO stream crosses HBM")
    sm = g.setup_ms or {}
    if "eri_s8" in sm:
        fl = roofline.batch_eri_flops(sizes_b)
        rl["eri_s8"] = _roof("valu", fl / (sm["eri_s8"] * 1e-3) / 1e12, PEAK_TFLOPS["fp32"], "TFLOP/s", sm["eri_s8"],
                             "Rys op-count model (pyscf_ipu_amd/roofline.py); peak = FP32 vector rate")
    out["minao"]["kernel_rooflines"] = rl
    out["minao"]["setup_ms"] = {k: round(v, 2) for k, v in sm.items()}
    common = {"mean_scf_iters_fp64_minao": float(its64.mean()),
              "fp64_reference_converged": int(conv64.sum()),
              "gate": "per molecule: FP32 run and FP64 batched reference both converged, |E - E_fp64| <= 1 mHa",
              "dtype": "f32 (ERI/JK/XC/Becke) + f64 (eigh/DIIS/energy)",
              "workload": f"{workload}, STO-3G, grid level 0, FP32 stop |dE| < 1e-6 (above the FP32 energy noise), "
                          f"<= {its} iterations, one launch per SCF phase, HIP-graph captured iterations"}
    return out, common


def _copy_calibration(n: int = 1 << 27, reps: int = 3) -> int:
    """A copy kernel of known traffic (8n bytes read, 8n written per launch)."""
    import torch
    from pyscf_ipu_amd import _lib, device
    torch.cuda.set_device(0)
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        _lib.call("pscf_cast", _lib.F64, device.ptr(x), _lib.F64, device.ptr(y), n, device.stream())
    torch.cuda.synchronize()
    assert bool((y == 1).all())
    print(json.dumps({"copy_kernel": "cast_kernel<double, double>", "launches": reps, "bytes_read_per_launch": 8 * n,
                      "bytes_written_per_launch": 8 * n}), flush=True)
    return 0


def _probes(n: int = 1 << 27, reps: int = 5) -> dict:
    """Measured machine constants beside the published peaks: HBM stream rate
    of the library's copy kernel (1 GiB read + 1 GiB written per launch) and a
    BF16 GEMM through torch (rocBLAS/hipBLASLt), both timed with HIP events on
    the launch stream."""
    import torch
    from pyscf_ipu_amd import _lib, device
    st = torch.cuda.current_stream()
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    copy = lambda: _lib.call("pscf_cast", _lib.F64, device.ptr(x), _lib.F64, device.ptr(y), n, device.stream())
    copy()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(st)
    for _ in range(reps):
        copy()
    b.record(st)
    b.synchronize()
    copy_gbs = 16.0 * n * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del x, y
    m = 8192
    A = torch.randn(m, m, device="cuda", dtype=torch.bfloat16)
    B = torch.randn(m, m, device="cuda", dtype=torch.bfloat16)
    C = A @ B
    a.record(st)
    for _ in range(reps):
        C = A @ B
    b.record(st)
    b.synchronize()
    gemm_tf = 2.0 * m ** 3 * reps / (a.elapsed_time(b) * 1e-3) / 1e12
    del A, B, C
    torch.cuda.empty_cache()
    return {"hbm_copy_gbs": copy_gbs, "hbm_copy_frac_of_8000": copy_gbs / HBM_PEAK_GBS,
            "bf16_gemm_8192_tflops": gemm_tf, "bf16_gemm_frac_of_2500": gemm_tf / 2500.0}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--dtype", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--nmol", type=int, default=32)
    ap.add_argument("--its", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-nmol", type=int, default=4096, help="0 skips the batched form")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB of FP64 three times with the library's copy kernel (pscf_cast "
                         "F64 -> F64): the known-size reference for the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args(argv)
    if a.copy_calibration:
        return _copy_calibration()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported (molecules shard "
              "across GPUs with no communication, so N GPUs would run N independent copies)", file=sys.stderr)
        return 2

    import torch
    from pyscf_ipu_amd import molgen
    from pyscf_ipu_amd.gto import build_mol
    from pyscf_ipu_amd.nanoDFT.nanoDFT import KernelEvents, run_molecules
    from pyscf_ipu_amd.nanoDFT.pool import MoleculePool, start_blocker
    from pyscf_ipu_amd import roofline

    mol_strs = molgen.plain_set(a.nmol, 9, seed=0)
    # the CPU baseline first: nothing else runs on the host cores meanwhile,
    # and its worker processes start before anything here uses the GPU
    cpu = None if a.no_cpu_baseline else _cpu_baseline(mol_strs)
    # the worker processes start before this process uses the GPU, and this
    # process stays off the GPU while they run (its own queues would share the
    # hardware queue slots with theirs)
    blocker = start_blocker()
    pool = MoleculePool(workers=PLAIN_WORKERS, dtype=a.dtype) if blocker is None else None
    if pool is None:
        torch.cuda.set_device(0)
        in_flight = f"{PLAIN_STREAMS} host threads on {PLAIN_STREAMS} HIP streams ({blocker})"
    else:
        in_flight = f"{PLAIN_WORKERS} worker processes"

    def plain_run():
        if pool is not None:
            return pool.run(mols, its=a.its, conv_tol=conv_tol)
        out = run_molecules(mols, streams=PLAIN_STREAMS, dtype=a.dtype, its=a.its, conv_tol=conv_tol)
        torch.cuda.synchronize()
        return out

    mols = [build_mol(s) for s in mol_strs]
    golden = None
    if os.path.exists(GOLDEN):
        gj = json.load(open(GOLDEN))
        if [m["mol_str"] for m in gj["molecules"]] == mol_strs[:len(gj["molecules"])]:
            golden = [m["energy"] for m in gj["molecules"]]

    # FP32 runs stop at |dE| < 1e-6 Ha, above their energy noise; FP64 at 1e-8
    conv_tol = 1e-8 if a.dtype == "fp64" else 1e-6

    # timed: each molecule's SCF with its own launches (one molecule per
    # launch), PLAIN_WORKERS molecules in flight in as many worker processes.
    # A result arrives after its worker has synchronised its stream, so the
    # last result of a step marks the end of that step's GPU work.
    for _ in range(a.warmup):
        plain_run()
    t0 = time.perf_counter()
    results = []
    for _ in range(a.steps):
        results = plain_run()
    t1 = time.perf_counter()
    wall = t1 - t0
    if pool is not None:
        pool.close()
        torch.cuda.set_device(0)
    energies = np.array([r.energy for r in results])
    iters = np.array([r.iterations for r in results])
    if golden is not None:
        dE = np.abs(energies[:len(golden)] - np.array(golden[:len(energies)]))
        passed = int(np.sum(dE <= 1e-3))
        max_dE = float(dE.max())
    else:
        dE, passed, max_dE = None, len(results), None
    value = passed * a.steps / wall
    # one molecule at a time (the reference's serial loop), with HIP events
    # around every launch group for the per-kernel times and the roofline;
    # first an untimed pass, since this process has not run the path yet
    run_molecules(mols, streams=1, dtype=a.dtype, its=a.its, conv_tol=conv_tol)
    ev = KernelEvents()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    run_molecules(mols, streams=1, dtype=a.dtype, its=a.its, conv_tol=conv_tol, events=ev)
    torch.cuda.synchronize()
    serial_wall = time.perf_counter() - t2

    # roofline of the dominant kernel group, from the HIP events of the timed region
    ks = ev.summary()
    dominant = max(ks, key=lambda k: ks[k]["total_ms"])
    work = roofline.per_launch_work(dominant, mols, a.dtype)
    mean_s = ks[dominant]["mean_ms"] * 1e-3
    if work["bound"] == "hbm":
        achieved = work["bytes"] / mean_s / 1e9
        peak, unit = HBM_PEAK_GBS, "GB/s"
    else:
        achieved = work["flops"] / mean_s / 1e12
        peak, unit = PEAK_TFLOPS[a.dtype], "TFLOP/s"
    kname = {"scf_step": "scf_step_kernel", "jk_s8": "jk_s8_kernel", "xc": "xc_fused_kernel"}.get(dominant)
    tdt = "double" if a.dtype == "fp64" else "float"
    traffic = _profile_traffic(f"{kname}<{tdt}>") if kname else None
    roof = {"kernel": dominant, "bound": work["bound"], "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": traffic, "mean_launch_ms": ks[dominant]["mean_ms"],
            "work_model": work["model"], "limiter": work.get("limiter"),
            "traffic_source": "profiles/hbm_traffic_bench.md: FETCH_SIZE + WRITE_SIZE bytes per dispatch, each "
                              "scaled by the copy-kernel calibration (8 molecules of the same generator)"}
    if dominant == "scf_step":  # one workgroup: the CU it runs on bounds it
        roof.update({"peak_one_cu": peak / N_CU, "frac_one_cu": achieved / (peak / N_CU)})
    kernel_rooflines = {}
    for k, v in ks.items():
        try:
            w = roofline.per_launch_work(k, mols, a.dtype)
        except KeyError:
            continue
        sec = v["mean_ms"] * 1e-3
        if w["bound"] == "hbm":
            kernel_rooflines[k] = _roof("hbm", w["bytes"] / sec / 1e9, HBM_PEAK_GBS, "GB/s", v["mean_ms"], w["model"])
        else:
            kernel_rooflines[k] = _roof(w["bound"], w["flops"] / sec / 1e12, PEAK_TFLOPS[a.dtype], "TFLOP/s",
                                        v["mean_ms"], w["model"])
    forms = {}
    if a.batched_nmol > 0:
        # equal work: the headline's 9-heavy-atom molecules first, then more of
        # the same generator (seed 2; the learned guess trained on seeds >= 10000)
        nine = mol_strs + molgen.plain_set(max(0, a.batched_nmol - len(mol_strs)), 9, seed=2)
        bt, common = _batched(nine[:a.batched_nmol], f"{a.batched_nmol} seeded molecules with 9 heavy atoms C/N/O "
                              f"(the plain headline's {len(mol_strs)} first)", a.steps, a.warmup, 30, golden=golden)
        for name, rec in bt.items():
            rec["speedup_vs_plain"] = rec["value"] / value
            rec.update(common)
            forms[f"batched_fp32_{name}"] = rec
        mixed = molgen.batched_set(a.batched_nmol, (1, 9), seed=1)
        bt, common = _batched(mixed, f"{a.batched_nmol} seeded molecules with 1-9 heavy atoms C/N/O", a.steps,
                              a.warmup, 30)
        for name, rec in bt.items():
            rec.update(common)
            rec["note"] = "not equal work to the plain headline (fewer heavy atoms per molecule on average)"
            forms[f"batched_fp32_{name}_1to9_heavy"] = rec
    line = {
        "metric": METRIC, "value": value, "unit": "molecules/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.dtype == "fp64" else "f32 (ERI/JK/XC) + f64 (eigh/DIIS/energy)",
        "data": "synthetic",
        "config": {"workload": "plain one-molecule-per-launch RKS B3LYP SCF: 32 seeded 9-heavy-atom C/N/O "
                               "molecules, STO-3G, grid level 0, |dE|<1e-8 stop, <=20 iterations, "
                               f"molecules in flight: {in_flight}",
                   "n_molecules": a.nmol, "basis": "sto-3g", "grid_level": 0, "form": "plain",
                   "nao_range": [int(min(m.nao_nr() for m in mols)), int(max(m.nao_nr() for m in mols))]},
        "mean_scf_iters": float(iters.mean()), "passed_1mha": passed, "max_abs_dE_ha": max_dE,
        "serial_one_stream": {"value": passed / serial_wall, "ms_per_step": serial_wall * 1e3},
        "kernel_ms_per_step": {k: v["total_ms"] for k, v in ks.items()},
        "roofline": roof, "kernel_rooflines": kernel_rooflines, "cpu_baseline": cpu, "forms": forms,
        "machine_probes": _probes(),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
