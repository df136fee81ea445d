"""Training of the learned-guess pair network on densities that the batched
GPU SCF itself converges.

    python -m pyscf_ipu_amd.train_guess --molecules 200000 --out pyscf_ipu_amd/weights/pair_mlp_sto3g.npz

Data: seeded generator molecules (1-9 heavy atoms C/N/O, hydrogens filling
the valences; seeds disjoint from the evaluation sets) are converged in
batches of 4096 by BatchRKS (FP32 ERI/JK/XC, FP64 eigh/DIIS); for every
ordered atom pair within 5 Angstrom of a converged molecule the features and
the 5x5 block of its converged density stay resident in HBM (BF16 features,
FP32 targets and masks).  Training: Adam with a cosine learning-rate decay on
random minibatches, every step being the library's BF16 MFMA forward and
backward kernels and its Adam kernel.
"""
from __future__ import annotations

import argparse
import math
import multiprocessing as mp
import os
import sys
import time
from typing import Dict, List, Sequence

import numpy as np
import torch

from . import _lib, batch, device, learned_guess as lg, molgen


def _gen_chunk(args):
    n, seed = args
    return molgen.batched_set(n, (1, 9), seed=seed)


def molecule_strings(n: int, seed0: int = 10_000, workers: int = 0) -> List[str]:
    """n generator molecules from seeds seed0, seed0+1, ... (1000 per seed)."""
    jobs = [(min(1000, n - k), seed0 + k // 1000) for k in range(0, n, 1000)]
    if workers > 1 and len(jobs) > 1:
        with mp.get_context("spawn").Pool(workers) as pool:
            parts = pool.map(_gen_chunk, jobs)
    else:
        parts = [_gen_chunk(j) for j in jobs]
    return [s for p in parts for s in p]


def targets(b: batch.BatchRKS, info: np.ndarray, valid: torch.Tensor):
    P = len(info)
    d_info = device.to_dev(info)
    Y = device.empty((max(P, 1), lg.NOUT_PAD), torch.float32)
    mask = device.empty((max(P, 1), lg.NOUT_PAD), torch.float32)
    _lib.call("pscf_pair_targets", P, device.ptr(d_info), device.ptr(valid), device.ptr(b.d_desc), device.ptr(b.D),
              device.ptr(Y), device.ptr(mask), device.stream())
    return Y[:P], mask[:P]


def generate(mol_strs: Sequence[str], dtype: str = "fp32", chunk: int = 4096, its: int = 30,
             conv_tol: float = 1e-7, log=None) -> Dict[str, torch.Tensor]:
    """Converge the molecules in batches and keep (features, targets, mask)
    of every pair of every converged molecule on the device."""
    X, Y, Mk = [], [], []
    n_conv = 0
    for c0 in range(0, len(mol_strs), chunk):
        t0 = time.perf_counter()
        mols = batch.parse(mol_strs[c0:c0 + chunk])
        b = batch.BatchRKS(mols, dtype=dtype)
        r = b.run(max_its=its, conv_tol=conv_tol)
        ok = np.where(r.converged)[0]
        n_conv += len(ok)
        pairs, info = lg.pair_lists(b, ok)
        X16, _, valid = lg.features(b, pairs)
        y, mk = targets(b, info, valid)
        X.append(X16[:len(pairs)])
        Y.append(y)
        Mk.append(mk)
        del b
        if log:
            torch.cuda.synchronize()
            log(f"data: molecules {c0 + len(mols)}/{len(mol_strs)} converged {n_conv} pairs "
                f"{sum(len(x) for x in X)} ({time.perf_counter() - t0:.1f} s)")
    data = {"X16": torch.cat(X), "Y": torch.cat(Y), "mask": torch.cat(Mk)}
    data["n_molecules"] = n_conv
    return data


def train(model: lg.PairMLP, data: Dict[str, torch.Tensor], steps: int, batch: int = 4096, lr: float = 1e-3,
          lr_min_frac: float = 0.02, seed: int = 0, log_every: int = 500, log=print) -> List[float]:
    """Adam with cosine decay on random minibatches; returns the loss of every
    step (masked mean squared error of the minibatch)."""
    X16, Y, mask = data["X16"], data["Y"], data["mask"]
    P = X16.shape[0]
    gen = torch.Generator(device=X16.device)
    gen.manual_seed(seed)
    # one normaliser for all minibatches: the expected number of valid
    # entries in a minibatch
    nvalid = batch * max(float(mask.sum().item()), 1.0) / P
    losses = []
    t0 = time.perf_counter()
    for s in range(steps):
        idx = torch.randint(0, P, (batch,), device=X16.device, generator=gen)
        xb, yb, mb = X16.index_select(0, idx), Y.index_select(0, idx), mask.index_select(0, idx)
        loss = model.loss_and_grads(xb, yb, mb, batch, nvalid=nvalid)
        frac = lr_min_frac + (1 - lr_min_frac) * 0.5 * (1 + math.cos(math.pi * s / max(steps - 1, 1)))
        model.adam(lr * frac)
        losses.append(loss.clone())
        if log_every and (s + 1) % log_every == 0:
            log(f"train: step {s + 1}/{steps} loss {float(loss.item()):.3e} ({time.perf_counter() - t0:.1f} s)")
    return [float(x) for x in torch.cat(losses).cpu().numpy()]


def evaluate(model: lg.PairMLP, data: Dict[str, torch.Tensor], batch: int = 1 << 16) -> float:
    """Masked MSE over a whole data set (FP64 accumulation)."""
    X16, Y, mask = data["X16"], data["Y"], data["mask"]
    tot, n = 0.0, 0.0
    for r0 in range(0, X16.shape[0], batch):
        B = min(batch, X16.shape[0] - r0)
        pred = model.forward(X16, B, r0)
        d = (pred - Y[r0:r0 + B]) * mask[r0:r0 + B]
        tot += float((d.double() ** 2).sum().item())
        n += float(mask[r0:r0 + B].sum().item())
    return tot / max(n, 1.0)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--molecules", type=int, default=200_000)
    ap.add_argument("--val-molecules", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=60_000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=lg.DEFAULT_WEIGHTS)
    ap.add_argument("--eval-molecules", type=int, default=1024)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    t0 = time.perf_counter()
    strs = molecule_strings(a.molecules, 10_000, a.workers)
    val = molecule_strings(a.val_molecules, 900_000, a.workers)
    log(f"generated {len(strs)} + {len(val)} molecules in {time.perf_counter() - t0:.1f} s")
    device.require_gpu()
    data = generate(strs, log=log)
    vdata = generate(val, log=log)
    model = lg.PairMLP(seed=0)
    hist = train(model, data, a.steps, a.batch, a.lr, log=log)
    v = evaluate(model, vdata)
    log(f"validation masked MSE {v:.3e}; final train loss {np.mean(hist[-100:]):.3e}")
    model.save(a.out, {"train_molecules": data["n_molecules"], "train_pairs": data["X16"].shape[0],
                       "steps": a.steps, "batch": a.batch, "lr": a.lr, "val_mse": v,
                       "train_loss_last100": float(np.mean(hist[-100:]))})
    log(f"saved {a.out} ({time.perf_counter() - t0:.1f} s total)")
    if a.eval_molecules:
        mols = batch.parse(molecule_strings(a.eval_molecules, 800_000, a.workers))
        base = lg.solve(mols, None, max_its=30)
        guess = lg.solve(mols, model, max_its=30)
        dE = np.abs(guess.result.energy - base.result.energy)
        log(f"held-out {len(mols)}: mean iterations minao {base.iterations_total.mean():.2f} learned "
            f"{guess.iterations_total.mean():.2f} (fallback {int(guess.fallback.sum())}); converged "
            f"{int(base.result.converged.sum())} / {int(guess.result.converged.sum())}; max |dE| {dE.max():.2e} Ha")
    return 0


if __name__ == "__main__":
    sys.exit(main())
