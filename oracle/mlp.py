"""ORACLE (test infrastructure only): FP64 numpy restatement of the learned
initial-guess network and its training step.

There is no counterpart in the reference (its guess is PySCF's MINAO,
pyscf_ipu/nanoDFT/nanoDFT.py:233); the network is specified by the task:
per ordered atom pair within 5 Angstrom, inputs = one-hot(Z_a) ++ one-hot(Z_b)
(H, C, N, O) ++ 32 Gaussian RBFs of the distance (centres 0..5 Angstrom,
sigma 0.15 Angstrom) ++ unit direction a->b (zero for a == b); MLP
43 -> 256 -> 256 -> 256 -> 25 with SiLU; output = the 5x5 STO-3G density
block D[AOs of a, AOs of b] (rows/cols beyond an atom's AO count masked).
Loss = mean over valid entries of (pred - target)^2.  Adam (beta1 0.9,
beta2 0.999, eps 1e-8).

Pins: finite differences of the loss for every parameter tensor
(tests/test_learned_guess.py), so the GPU gradients are checked against
derivatives that do not share code with either implementation.
"""
from __future__ import annotations

import numpy as np

ELEM_INDEX = {1: 0, 6: 1, 7: 2, 8: 3}
NRBF, RCUT_A, SIGMA_A = 32, 5.0, 0.15
BOHR = 0.52917721092
NIN, NHID, NOUT = 43, 256, 25
AO_PER = {1: 1, 6: 5, 7: 5, 8: 5}


def pair_features(Z: np.ndarray, R_bohr: np.ndarray):
    """Ordered pairs (a, b) with |R_a - R_b| < 5 Angstrom (a == b included):
    returns (pairs (P, 2), features (P, 43), mask (P, 25))."""
    n = len(Z)
    R = R_bohr * BOHR
    d = np.linalg.norm(R[:, None] - R[None], axis=-1)
    a, b = np.nonzero(d < RCUT_A)
    P = len(a)
    X = np.zeros((P, NIN))
    X[np.arange(P), [ELEM_INDEX[int(z)] for z in Z[a]]] = 1.0
    X[np.arange(P), 4 + np.array([ELEM_INDEX[int(z)] for z in Z[b]])] = 1.0
    mu = np.linspace(0.0, RCUT_A, NRBF)
    X[:, 8:8 + NRBF] = np.exp(-((d[a, b][:, None] - mu[None]) ** 2) / (2 * SIGMA_A ** 2))
    v = R[b] - R[a]
    dist = d[a, b]
    nz = dist > 0
    X[nz, 8 + NRBF:] = v[nz] / dist[nz, None]
    mask = np.zeros((P, 5, 5))
    na = np.array([AO_PER[int(z)] for z in Z[a]])
    nb = np.array([AO_PER[int(z)] for z in Z[b]])
    for k in range(P):
        mask[k, :na[k], :nb[k]] = 1.0
    return np.stack([a, b], axis=1), X, mask.reshape(P, 25)


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def init_params(seed: int = 0):
    rng = np.random.default_rng(seed)
    dims = [NIN, NHID, NHID, NHID, NOUT]
    params = []
    for i in range(4):
        w = rng.normal(size=(dims[i], dims[i + 1])) * np.sqrt(2.0 / dims[i])
        params += [w, np.zeros(dims[i + 1])]
    return params


def forward(params, X):
    acts = [X]
    pre = []
    h = X
    for i in range(4):
        z = h @ params[2 * i] + params[2 * i + 1]
        pre.append(z)
        h = silu(z) if i < 3 else z
        acts.append(h)
    return h, (acts, pre)


def loss_and_grads(params, X, Y, mask):
    out, (acts, pre) = forward(params, X)
    nvalid = max(mask.sum(), 1.0)
    diff = (out - Y) * mask
    loss = float(np.sum(diff * diff) / nvalid)
    g = 2.0 * diff / nvalid
    grads = [None] * 8
    for i in range(3, -1, -1):
        if i < 3:
            g = g * dsilu(pre[i])
        grads[2 * i] = acts[i].T @ g
        grads[2 * i + 1] = g.sum(axis=0)
        g = g @ params[2 * i].T
    return loss, grads


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    out = []
    for k, (p, gr) in enumerate(zip(params, grads)):
        m[k] = b1 * m[k] + (1 - b1) * gr
        v[k] = b2 * v[k] + (1 - b2) * gr * gr
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        out.append(p - lr * mh / (np.sqrt(vh) + eps))
    return out
