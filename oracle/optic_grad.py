"""Float64 torch restatement of the differentiable optic lobe (TEST INFRASTRUCTURE ONLY: the checker of
flyhip.adjoint, never imported by the product).

The forward of flyverse/optic.py:525-551 (no spiking feedback) written functionally like
OpticLobe._step_frame_grad (optic.py:606-676), with per-type tau / operating point and per-pair gains as the
parameters, in float64; torch autograd gives the reference gradient.  Pinned against central finite differences in
float64 (tests/test_adjoint.py), which fixes it independently of the adjoint's derivation.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch
from torch.utils.checkpoint import checkpoint as _checkpoint


def _step(v, adapt, w, a, b, P_t, rows, cols, g_rr, g_ad, a_ad):
    """One step of the rate dynamics (flyverse/optic.py:528-551): returns (v', adapt')."""
    dr = (v + b).clamp(0.0, 1.0) - b
    y = torch.zeros_like(v).index_add(0, rows, w[:, None] * dr[cols])
    inp = g_rr * y + P_t - g_ad * adapt
    return inp + (v - inp) * a, dr + (adapt - dr) * a_ad


def loss64(tau, bias, gain, W0: sp.csr_matrix, unit_type, P, target, readout, dt=5.0, g_rr=1.0, g_ad=1.0,
           adapt_tau=400.0, per_step_checkpoint=False):
    """tau, bias (K,), gain (K, K) float64 tensors (may require grad); P (T, n, B); target (T, R, B).

    per_step_checkpoint: keep only (v, adapt) per step for the backward and recompute each step's product there
    (torch.utils.checkpoint), so the full-size problem fits in host memory; the arithmetic is the same."""
    W0 = sp.coo_matrix(W0)
    rows = torch.as_tensor(W0.row.astype(np.int64))
    cols = torch.as_tensor(W0.col.astype(np.int64))
    w0 = torch.as_tensor(W0.data.astype(np.float64))
    ut = torch.as_tensor(np.asarray(unit_type, np.int64))
    K = tau.shape[0]
    w = gain.reshape(-1)[ut[rows] * K + ut[cols]] * w0
    a = torch.exp(-dt / tau)[ut][:, None]
    b = bias[ut][:, None]
    a_ad = float(np.exp(-dt / adapt_tau))
    ro = torch.as_tensor(np.asarray(readout, np.int64))
    T, n, B = P.shape
    v = torch.zeros(n, B, dtype=torch.float64)
    adapt = torch.zeros(n, B, dtype=torch.float64)
    total = torch.zeros((), dtype=torch.float64)
    for t in range(T):
        P_t = torch.as_tensor(np.asarray(P[t]), dtype=torch.float64)
        args = (v, adapt, w, a, b, P_t, rows, cols, g_rr, g_ad, a_ad)
        if per_step_checkpoint and torch.is_grad_enabled():
            v, adapt = _checkpoint(_step, *args, use_reentrant=False)
        else:
            v, adapt = _step(*args)
        r = (v + b).clamp(0.0, 1.0)
        tgt = torch.as_tensor(np.asarray(target[t]), dtype=torch.float64)
        total = total + ((r[ro] - tgt) ** 2).sum()
    return total / (T * B * len(readout))
