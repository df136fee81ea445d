"""Learned initial guess: oracle pins and host logic (CPU).

The network has no counterpart in the reference, so the oracle
(oracle/mlp.py) is pinned by central finite differences of its loss for
every parameter tensor, and its features by closed forms (one-hot rows,
the RBF peak at the centre nearest the distance, unit vectors)."""
import numpy as np
import pytest

from oracle import mlp as omlp
from pyscf_ipu_amd import batch, learned_guess as lg, molgen


def _tiny_problem(seed=0, P=7):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(P, omlp.NIN))
    Y = rng.normal(size=(P, omlp.NOUT))
    mask = (rng.random((P, omlp.NOUT)) < 0.8).astype(float)
    return X, Y, mask


def test_oracle_gradients_match_finite_differences():
    params = omlp.init_params(3)
    # small random biases so every bias gradient path is exercised
    rng = np.random.default_rng(4)
    for k in range(1, 8, 2):
        params[k] = rng.normal(size=params[k].shape) * 0.1
    X, Y, mask = _tiny_problem()
    loss, grads = omlp.loss_and_grads(params, X, Y, mask)
    h = 1e-6
    for k, p in enumerate(params):
        flat = p.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp, _ = omlp.loss_and_grads(params, X, Y, mask)
            flat[idx] = old - h
            lm, _ = omlp.loss_and_grads(params, X, Y, mask)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[k].reshape(-1)[idx]) < 1e-6 * max(1.0, abs(fd)), (k, idx, fd)


def test_oracle_adam_first_step_is_signed_lr():
    # at t = 1 the bias-corrected update is -lr * g / (|g| + eps)
    params = [np.array([1.0, -2.0, 0.5])]
    g = [np.array([0.3, -4.0, 1e-3])]
    out = omlp.adam_step(params, g, [np.zeros(3)], [np.zeros(3)], 1, lr=0.01)
    np.testing.assert_allclose(out[0], params[0] - 0.01 * g[0] / (np.abs(g[0]) + 1e-8), rtol=1e-12)


def test_oracle_features_closed_forms():
    Z = np.array([6, 1, 8])
    R = np.array([[0.0, 0.0, 0.0], [1.09, 0.0, 0.0], [0.0, 0.0, 6.0]]) / omlp.BOHR
    pairs, X, mask = omlp.pair_features(Z, R)
    got = {tuple(p) for p in pairs.tolist()}
    assert got == {(0, 0), (1, 1), (2, 2), (0, 1), (1, 0)}  # O is 6 A from C, 6.1 A from H
    k = [tuple(p) for p in pairs.tolist()].index((0, 1))
    assert X[k, 1] == 1 and X[k, 4] == 1 and X[k, :8].sum() == 2  # C then H
    rbf = X[k, 8:40]
    assert np.argmax(rbf) == int(round(1.09 / (5.0 / 31)))
    np.testing.assert_allclose(X[k, 40:], [1, 0, 0])
    assert mask[k].sum() == 5  # 5 AOs of C x 1 AO of H


def test_rerun_mask_selects_unconverged_and_small_gap():
    """Re-run from MINAO: not converged (unless failed), or converged with a
    gap below the threshold; a converged molecule at or above it is kept."""
    conv = np.array([True, True, False, False, True, True])
    failed = np.array([False, False, False, True, False, True])
    gap = np.array([0.20, 0.03, 0.01, 0.30, lg.GAP_RECHECK, 0.01])
    redo, small_gap = lg.rerun_mask(conv, failed, gap)
    assert redo.tolist() == [False, True, True, False, False, False]
    assert small_gap.tolist() == [False, True, False, False, False, False]


def test_flat_layout_round_trip():
    flat = lg.init_params(1)
    parts = lg.unflatten(flat)
    shapes = [p.shape for p in parts]
    assert shapes == [(43, 256), (256,), (256, 256), (256,), (256, 256), (256,), (256, 25), (25,)]
    # padding is zero
    off, r, c = lg.OFFSETS["W0"]
    assert np.all(flat[off + 43 * c: off + r * c] == 0)
    off, r, c = lg.OFFSETS["W3"]
    assert np.all(flat[off:off + r * c].reshape(r, c)[:, 25:] == 0)


def test_pair_lists_match_oracle():
    strs = molgen.batched_set(12, (1, 5), seed=8)
    mols = batch.parse(strs)
    b = batch.BatchRKS(mols)
    pairs, info = lg.pair_lists(b)
    atm0 = np.cumsum(b.natm) - b.natm
    for m, (Z, R) in enumerate(mols):
        sel = info[:, 0] == m
        mine = {(a - atm0[m], c - atm0[m]) for a, c in pairs[sel].tolist()}
        op, _, omask = omlp.pair_features(Z, R)
        assert mine == {tuple(p) for p in op.tolist()}
        # AO counts per atom and AO offsets consistent with the basis
        for (a, na_, c, nc_) in info[sel][:, 1:].tolist():
            assert na_ in (1, 5) and nc_ in (1, 5)
            assert 0 <= a and a + na_ <= b.nao[m] and c + nc_ <= b.nao[m]

