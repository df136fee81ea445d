"""Molecules in flight across GPU worker processes (the plain form's pool).

The plain form runs every molecule with its own launches (`RKS(mol).run()`,
one molecule per launch as nanoDFT runs them,
pyscf_ipu/nanoDFT/nanoDFT.py:259-336).  A molecule's SCF is latency-bound,
so throughput comes from molecules in flight.  Host threads in one process
(`run_molecules`) share one interpreter lock, and each molecule's Python
set-up (grid, descriptors, device buffers, ~15 library calls) holds it; the
threads' queues then sit idle for a large part of the run.  Worker processes
each own an interpreter and a GPU context, so their set-ups proceed at once.

Workers are started (spawn: fork + exec of a fresh interpreter) by a process
that has not used the GPU yet; `MoleculePool` refuses to start otherwise.
Tasks go through one shared queue, so the workers balance themselves.
"""
from __future__ import annotations

import multiprocessing as mp
import os
import queue as _queue
import traceback
from typing import List, Optional, Sequence


def start_blocker() -> Optional[str]:
    """Why this process must not start worker processes, or None.  Workers are
    started by fork + exec, which must not follow GPU initialisation here: by
    torch, or by a profiler's preloaded library (it initialises the GPU before
    the program starts)."""
    import torch
    if torch.cuda.is_initialized():
        return "this process already uses the GPU"
    if any(k.startswith("ROCPROF") for k in os.environ) or "rocprofiler" in os.environ.get("LD_PRELOAD", ""):
        return "a profiler is attached to this process"
    return None


def _worker(task_q, result_q, level: int, dtype: str):
    try:
        import torch
        from pyscf_ipu_amd.nanoDFT.nanoDFT import RKS
        torch.cuda.set_device(0)
        torch.zeros(1, device="cuda")  # context created before the first task
        result_q.put(("ready", None, None))
    except BaseException:
        result_q.put(("error", -1, traceback.format_exc()))
        return
    while True:
        task = task_q.get()
        if task is None:
            break
        idx, mol, its, conv_tol, diis = task
        try:
            res = RKS(mol, level=level, dtype=dtype).run(its=its, conv_tol=conv_tol, diis=diis)
            result_q.put(("ok", idx, res))
        except BaseException:
            result_q.put(("error", idx, traceback.format_exc()))


class MoleculePool:
    """`workers` GPU processes, each running one molecule at a time with its own
    launches.  `run(mols)` returns the `SCFResult`s in input order."""

    def __init__(self, workers: int = 8, level: int = 0, dtype: str = "fp64", start_timeout: float = 600.0):
        reason = start_blocker()
        if reason is not None:
            raise RuntimeError(f"MoleculePool: {reason}; create the pool before this process uses the GPU "
                               "(its workers are started by fork + exec)")
        if workers < 1:
            raise ValueError("MoleculePool: workers must be >= 1")
        ctx = mp.get_context("spawn")
        self.workers = workers
        self._tasks = ctx.Queue()
        self._results = ctx.Queue()
        self._procs = [ctx.Process(target=_worker, args=(self._tasks, self._results, level, dtype), daemon=True)
                       for _ in range(workers)]
        for p in self._procs:
            p.start()
        for _ in range(workers):
            kind, _, payload = self._next(start_timeout)
            if kind != "ready":
                self.close()
                raise RuntimeError(f"MoleculePool: a worker failed to start:\n{payload}")

    def _next(self, timeout: Optional[float]):
        """The next message of any worker; raises as soon as a worker process
        has exited (a crash, or a failed start) or the time limit passes."""
        waited = 0.0
        while True:
            try:
                return self._results.get(timeout=1.0)
            except _queue.Empty:
                waited += 1.0
                dead = [p.exitcode for p in self._procs if not p.is_alive()]
                if dead:
                    try:  # a message the worker sent before it exited
                        return self._results.get(timeout=1.0)
                    except _queue.Empty:
                        pass
                    raise RuntimeError(f"MoleculePool: a worker process exited (exit codes {dead})") from None
                if timeout is not None and waited >= timeout:
                    raise RuntimeError("MoleculePool: no message from the workers within the time limit") from None

    def run(self, mols: Sequence, its: int = 20, conv_tol: float = 1e-8, diis: bool = True,
            timeout: Optional[float] = 600.0) -> List:
        """Raises on the first failed molecule, dead worker or time-out, and the
        pool is closed then: the batch's other molecules are still in flight, and
        their results must not reach a later `run`."""
        if not self._procs:
            raise RuntimeError("MoleculePool: the pool is closed")
        for i, m in enumerate(mols):
            self._tasks.put((i, m, its, conv_tol, diis))
        out: List = [None] * len(mols)
        try:
            for _ in range(len(mols)):
                kind, idx, payload = self._next(timeout)
                if kind != "ok":
                    raise RuntimeError(f"MoleculePool: molecule {idx} failed in a worker:\n{payload}")
                out[idx] = payload
        except BaseException:
            self.close()
            raise
        return out

    def close(self):
        """Stops the workers: each finishes its molecule, then exits; one still
        running after a minute is terminated."""
        for p in self._procs:
            if p.is_alive():
                self._tasks.put(None)
        for p in self._procs:
            p.join(timeout=60)
            if p.is_alive():
                p.terminate()
                p.join(timeout=10)
        self._procs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
