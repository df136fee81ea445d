"""ORACLE (test infrastructure only): B3LYP energy density and derivatives.

Restates the reference's functional (pyscf_ipu/exchange_correlation/b3lyp.py:13-39,
lda.py:5, b88.py:13-29, lyp.py:13-45, vwn.py:13-54; formulas ported there from
XCFun) in torch FP64 on the CPU.  Derivatives come from torch.autograd, as the
reference takes them from jax.value_and_grad.

  b3lyp(rho (4,g), eps) -> (e_xc per particle, v_rho = dE/drho, v_norms = dE/d|grad rho|^2)

rho[1:4] is half the density gradient, as the reference's exchange_correlation
produces it (nanoDFT.py:109-112); it is doubled here like b3lyp.py:15.

Pins (tests/test_xc_pins.py): the LDA part against the closed-form Slater
constant (3/4)(3/pi)^(1/3); B88 at zero gradient against spin-resolved Slater
exchange and its gradient correction against Becke's published form in spin
variables; VWN5 against the Perdew-Zunger fit of the same Ceperley-Alder data
(within 0.5 mHa for r_s >= 1) and its high-density slope (1 - ln 2)/pi^2; LYP at
zero gradient against its closed form; the mixing against a0 = 0.2, ax = 0.72,
ac = 0.81; v_rho and v_norms of the whole and of each piece against central
finite differences.  The GPU kernel's independently hand-derived derivatives
are compared with this module (tests/test_boundary.py, tests/test_gpu_kernels.py).
Pointwise agreement with XCFun/PySCF themselves is not checked (neither is
installed here).
"""
from __future__ import annotations

import math

import numpy as np
import torch

CLIP_RHO_MIN = 1e-9
CLIP_RHO_MAX = 1e12


def _lda(rho):
    return -torch.exp(1 / 3 * torch.log(rho) - 0.30305460484554375)


def _b88(a, gaa):
    c1 = 4.0 / 3.0
    c2 = -8.0 / 3.0
    c3 = (-3.0 / 4.0) * (6.0 / math.pi) ** (1.0 / 3.0) * 2
    d = 0.0042
    log_a = torch.log(a / 2)
    na43 = torch.exp(log_a * c1)
    chi2 = gaa / 4 * torch.exp(log_a * c2)
    chi = torch.exp(torch.log(chi2) / 2)
    b88 = -(d * na43 * chi2) / (1.0 + 6 * d * chi * torch.arcsinh(chi)) * 2
    return c3 * na43 + b88


def _lyp(n, gnn):
    A, B, C, Dd = 0.04918, 0.132, 0.2533, 0.349
    CF = 0.3 * (3.0 * math.pi * math.pi) ** (2.0 / 3.0)
    c0 = 2.0 ** (11.0 / 3.0) * (1 / 2) ** (8 / 3)
    c1 = (1 / 3 + 1 / 8) * 4
    log_n = torch.log(n)
    icbrtn = torch.exp(log_n * (-1.0 / 3.0))
    P = 1.0 / (1.0 + Dd * icbrtn)
    omega = torch.exp(-C * icbrtn) * P
    delta = icbrtn * (C + Dd * P)
    n53 = torch.exp(log_n * (-5 / 3))
    return -A * (n * P + B * omega * 1 / 4 * (2 * CF * n * c0 + gnn * (60 - 14.0 * delta) / 36 * n53 - gnn * c1 * n53))


def _vwn(n):
    p = np.array([-0.10498, 0.0621813817393097900698817274255, 3.72744, 12.9352])
    f = p[0] * p[2] / (p[0] * p[0] + p[0] * p[2] + p[3]) - 1.0
    f_inv_p1 = 1 / f + 1
    sq = np.sqrt(4.0 * p[3] - p[2] * p[2])
    pre = p[2] * (1.0 / sq - p[0] / ((p[0] * p[0] + p[0] * p[2] + p[3]) * sq / (p[2] + 2.0 * p[0])))
    log_s_c = np.log(3.0 / (4 * np.pi)) / 6
    log_s = -torch.log(n) / 6 + log_s_c
    s = torch.exp(log_s)
    s2 = torch.exp(log_s * 2)
    z = sq / (2.0 * s + p[2])
    return n * p[1] * (log_s + f * torch.log(torch.sqrt(s2 + p[2] * s + p[3]) / (s - p[0]) ** f_inv_p1)
                       + pre * torch.arctan(z))


def b3lyp(rho: np.ndarray, eps: float = 0.0):
    rho = torch.as_tensor(np.asarray(rho, dtype=np.float64))
    r0 = torch.clamp(rho[0], CLIP_RHO_MIN, CLIP_RHO_MAX).detach().requires_grad_(True)
    grad = rho[1:4] * 2
    norms = (torch.sum(grad * grad, dim=0) + eps).detach().requires_grad_(True)
    e_lda = _lda(r0) * 0.08
    e_vwn = _vwn(r0) * 0.19
    e_b88 = _b88(r0, norms) * 0.72
    e_lyp = _lyp(r0, norms) * 0.81
    (d_lda,) = torch.autograd.grad(e_lda.sum(), r0, retain_graph=True)
    (d_vwn,) = torch.autograd.grad(e_vwn.sum(), r0, retain_graph=True)
    d_b88 = torch.autograd.grad(e_b88.sum(), (r0, norms), retain_graph=True)
    d_lyp = torch.autograd.grad(e_lyp.sum(), (r0, norms))
    e_xc = e_lda + (e_vwn + e_b88 + e_lyp) / r0
    v_rho = d_lda * 4 * r0 + d_vwn + d_b88[0] + d_lyp[0]
    v_norms = d_b88[1] + d_lyp[1]
    return e_xc.detach().numpy(), v_rho.detach().numpy(), v_norms.detach().numpy()


def exchange_correlation(dm: np.ndarray, ao: np.ndarray, weights: np.ndarray, eps: float = 1e-20):
    """E_xc and V_xc as pyscf_ipu/nanoDFT/nanoDFT.py:105-123, FP64."""
    ao_dm = ao[0] @ dm                                    # (g, N)
    rho = np.einsum("gn,cgn->cg", ao_dm, ao)              # (4, g); rho[1:4] = grad rho / 2
    e_xc, v_rho, v_norms = b3lyp(rho, eps)
    E_xc = float(np.sum(rho[0] * weights * e_xc))
    v = np.concatenate([v_rho[None] / 2, 4 * v_norms[None] * rho[1:4]], axis=0) * weights
    V = ao[0].T @ np.einsum("cgn,cg->gn", ao, v)
    return E_xc, V + V.T, rho
