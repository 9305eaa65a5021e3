"""Seeded synthetic workload of the differentiable optic lobe: an 89,390-unit rate network with ~6 M signed
edges, 65 cell types, 32 drifting-grating movies of 200 steps at dt 5 ms and 734 readout units.

Build choices (not given by the workload statement), fixed here:
  * recurrent edges: lognormal in-degree (sigma 1) around nnz/n, uniform presynaptic cells, power-law synapse
    counts as flyhip.synth, sign fixed per presynaptic cell (70 % positive), each row normalised by the L2 norm of
    its counts (OpticParams.norm = "l2", flyverse/optic.py:245-248);
  * photoreceptors: 5,895 at uniform random (azimuth, elevation) in [-90, 90] x [-60, 60] degrees, 0.3 M
    photoreceptor->unit edges normalised the same way; gain_in = 3;
  * movies: intensity 0.5 + 0.5 sin(2 pi (x cos th + y sin th - speed t) / 30 deg), th ~ U[0, 2 pi),
    speed ~ U[0, 100] deg/s; the photoreceptor stage of flyverse/optic.py:496-517 at frame = dt turns it into the
    contrast a_pr, and P_t = gain_in * W_rp a_pr(t) is the held drive of step t;
  * parameters: tau_k ~ 20 ms * exp(N(0, 0.3)), b_k ~ N(0, 0.1), gain 1; targets come from a teacher run with
    tau * exp(N(0, 0.2)), b + N(0, 0.05), gain * exp(N(0, 0.1)).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from . import synth

N_UNITS, NNZ, N_TYPES, N_PR, NNZ_PR, T_STEPS, N_READOUT, MOVIES, DT = 89_390, 6_000_000, 65, 5_895, 300_000, 200, 734, 32, 5.0


def _normalised(rng, rows_n, cols_n, nnz):
    r, c = synth._block(rng, np.arange(rows_n), np.arange(cols_n), nnz)
    key = np.unique(r * cols_n + c)
    post, pre = key // cols_n, key % cols_n
    sign = np.where(rng.random(cols_n) < 0.7, 1.0, -1.0)
    counts = synth.power_law_counts(rng, len(key)).astype(np.float64)
    l2 = np.sqrt(np.bincount(post, weights=counts ** 2, minlength=rows_n))
    data = (counts * sign[pre] / np.maximum(l2, 1.0)[post]).astype(np.float32)
    M = sp.csr_matrix((data, (post, pre)), shape=(rows_n, cols_n))
    M.sum_duplicates()
    M.sort_indices()
    return M


def optic_network(seed=4, scale=1.0):
    """(W0 (n, n) csr float32, unit_type (n,), W_rp (n, n_pr) csr, photoreceptor positions (n_pr, 2) degrees)."""
    rng = np.random.default_rng(seed)
    n = max(8, int(round(N_UNITS * scale)))
    n_pr = max(4, int(round(N_PR * scale)))
    W0 = _normalised(rng, n, n, int(NNZ * scale))
    unit_type = rng.integers(0, N_TYPES, n)
    W_rp = _normalised(rng, n, n_pr, int(NNZ_PR * scale))
    pos = np.stack([rng.uniform(-90, 90, n_pr), rng.uniform(-60, 60, n_pr)], axis=1)
    return W0, unit_type, W_rp, pos


def grating_intensity(pos, T, movies=MOVIES, dt=DT, seed=5, wavelength=30.0):
    """(movies, T, n_pr) intensities of drifting sinusoidal gratings."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * math.pi, movies)
    speed = rng.uniform(0, 100, movies)
    t = np.arange(T) * dt / 1000.0
    proj = pos[:, 0][None] * np.cos(th)[:, None] + pos[:, 1][None] * np.sin(th)[:, None]     # (movies, n_pr)
    phase = 2 * math.pi * (proj[:, None, :] - speed[:, None, None] * t[None, :, None]) / wavelength
    return (0.5 + 0.5 * np.sin(phase)).astype(np.float32)


def photoreceptor_drive(intensity, W_rp, dt=DT, gain_in=3.0, tau_lp=10.0, tau_adapt=300.0, eps=0.02, clip=2.0):
    """P (T, n, movies): the photoreceptor stage (flyverse/optic.py:505-513, one cell per photoreceptor, fresh at
    t = 0) then gain_in * W_rp a_pr, in float64 rounded to float32 (input data of the workload)."""
    movies, T, n_pr = intensity.shape
    a_lp, a_ad = math.exp(-dt / tau_lp), math.exp(-dt / tau_adapt)
    I_lp = intensity[:, 0].astype(np.float64)
    I_mean = I_lp.copy()
    W = W_rp.astype(np.float64)
    P = np.empty((T, W_rp.shape[0], movies), np.float32)
    for t in range(T):
        I = intensity[:, t].astype(np.float64)
        I_lp = a_lp * I_lp + (1 - a_lp) * I
        I_mean = a_ad * I_mean + (1 - a_ad) * I_lp
        c = np.clip((I_lp - I_mean) / (I_mean + eps), -1.0, clip)
        P[t] = (gain_in * (W @ c.T)).astype(np.float32)
    return P


def initial_params(seed=6, K=N_TYPES):
    rng = np.random.default_rng(seed)
    tau = (20.0 * np.exp(rng.normal(0, 0.3, K))).astype(np.float32)
    bias = rng.normal(0, 0.1, K).astype(np.float32)
    gain = np.ones((K, K), np.float32)
    return tau, bias, gain


def teacher_params(tau, bias, gain, seed=7):
    rng = np.random.default_rng(seed)
    K = len(tau)
    return ((tau * np.exp(rng.normal(0, 0.2, K))).astype(np.float32),
            (bias + rng.normal(0, 0.05, K)).astype(np.float32),
            (gain * np.exp(rng.normal(0, 0.1, (K, K)))).astype(np.float32))
