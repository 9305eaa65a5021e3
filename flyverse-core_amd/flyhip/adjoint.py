"""Differentiable optic lobe with a hand-written adjoint (an addition: the reference differentiates its optic lobe
only through torch autograd, flyverse/optic.py:606-676, and has no adjoint or checkpointing).

The rate network of flyverse/optic.py:525-551 without spiking feedback, on 32 movies at once, with per-type time
constants tau, per-type operating points (biases) b and a gain per (postsynaptic type, presynaptic type) pair on the
normalised recurrent weights (the parameters OpticParams.tau_by_type / baseline_by_type / pair_gain expose,
optic.py:82-87, 251-267, 343-347):

    dr_t = clip(v_t + b, 0, 1) - b;  y_t = (G o W0) dr_t;  inp_t = g_rr y_t + P_t - g_ad adapt_t
    v_t+1 = inp_t + (v_t - inp_t) a,  a = exp(-dt / tau);  adapt_t+1 = dr_t + (adapt_t - dr_t) a_ad
    loss = mean over t = 1..T, movies and readout units of (clip(v_t + b, 0, 1) - target_t)^2

`OpticAdjoint.loss(tau, bias, gain)` runs the forward (FP32, csrc/adjoint.hip) keeping (v, adapt) every
`checkpoint` steps and every state of the last segment; `gradients()` walks the segments backwards, recomputing each
earlier segment's states from its checkpoint (with checkpoint >= T nothing is recomputed: the trajectory, about
46 MB per step at full scale, stays in HBM) and running the adjoint step by step: the transpose product (W^T applied to g_rr lambda_inp through the
transposed CSR), the clamp gates (gradient passes where 0 <= v + b <= 1, torch's convention for clamp), and a
per-edge accumulation of sum_b lambda_y dr for the pair gains.  Per-type and per-pair sums are fixed-order segmented
reductions, so the gradient is deterministic.  `OpticLossFn` wraps both in a torch.autograd.Function.
"""
from __future__ import annotations

import ctypes as ct
import math
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import torch

from . import hip

NL = 32
_P, _I, _F, _S, _L = ct.POINTER(ct.c_void_p), ct.c_int, ct.c_float, ct.c_void_p, ct.c_int64
_FP = ct.POINTER(ct.c_float)
SIGS = {
    "launch_adj_params": [_P, _I, _F, _S],
    "launch_adj_weights": [_P, _L, _S],
    "launch_adj_dr": [_P, _I, _S],
    "launch_adj_forward": [_P, _I, _FP, _I, _S],
    "launch_adj_back_ly": [_P, _I, _F, _S],
    "launch_adj_back_rows": [_P, _I, _FP, _I, _S],
    "launch_segment_sum": [_P, _I, _S],
    "launch_loss_total": [_P, _L, _F, _S],
}


def _lib(device):
    L = hip.lib(device)
    if not getattr(L, "_adj_sigs", False):
        for name, sig in SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = sig
            fn.restype = _I
        L._adj_sigs = True
    return L


def _segments(keys: np.ndarray, nseg: int):
    perm = np.argsort(keys, kind="stable").astype(np.int32)
    seg = np.zeros(nseg + 1, np.int64)
    np.cumsum(np.bincount(keys, minlength=nseg), out=seg[1:])
    return perm, seg.astype(np.int32)


@dataclass
class OpticConstants:
    dt: float = 5.0
    gain_rr: float = 1.0
    adapt_gain: float = 1.0
    adapt_tau: float = 400.0


class OpticAdjoint:
    """The C4-style training problem on the GPU.  W0: (n, n) normalised signed weights (post, pre); unit_type (n,)
    in [0, K); P: (T, n, 32) held input drive per step; target: (T, R, 32) for steps 1..T; readout: (R,) units.
    checkpoint: steps per recomputed segment; None keeps the whole trajectory (4 * T * n * 128 bytes) and recomputes
    nothing."""

    def __init__(self, W0: sp.spmatrix, unit_type, n_types: int, P, target, readout, consts=None, checkpoint=None,
                 device="cuda", grid=None):
        self.device = dev = torch.device(device)
        if not hip.is_rocm_gpu(dev):
            raise ValueError("the optic adjoint runs on a ROCm GPU only")
        _lib(dev)
        c = self.c = consts or OpticConstants()
        W0 = sp.csr_matrix(W0, dtype=np.float32)
        W0.sum_duplicates()
        W0.sort_indices()
        self.n = n = W0.shape[0]
        self.K = K = int(n_types)
        ut = np.asarray(unit_type, np.int64)
        self.T = int(P.shape[0])
        if tuple(P.shape) != (self.T, n, NL):
            raise ValueError(f"P must be (T, n, {NL})")
        self.R = len(readout)
        if tuple(target.shape) != (self.T, self.R, NL):
            raise ValueError(f"target must be (T, R, {NL})")
        self.checkpoint = self.T if checkpoint is None else int(checkpoint)
        if self.checkpoint < 1:
            raise ValueError("checkpoint must be >= 1")
        t = lambda a, dt=None: torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dt)
        rows = np.repeat(np.arange(n), np.diff(W0.indptr))
        bins = (ut[rows] * K + ut[W0.indices]).astype(np.int32)
        self.sched = t(np.argsort(-np.diff(W0.indptr), kind="stable").astype(np.int32))
        # the forward product's chunk-padded copy (hip.chunk_padded): weights of pads are 0 * gain
        fptr, fidx, fw0, _ = hip.chunk_padded(W0)
        fbin = np.zeros(len(fidx), np.int32)
        fbin[fidx >= 0] = bins
        self.fptr, self.fidx, self.fw0, self.fbin = t(fptr.astype(np.int32)), t(fidx), t(fw0), t(fbin)
        WT = W0.T.tocsr()
        WT.sort_indices()
        rowsT = np.repeat(np.arange(n), np.diff(WT.indptr))
        binsT = (ut[WT.indices] * K + ut[rowsT]).astype(np.int32)
        # the transposed product's chunk-padded copy
        bptr, bidx, bw0, _ = hip.chunk_padded(WT)
        bbin = np.zeros(len(bidx), np.int32)
        bbin[bidx >= 0] = binsT
        self.bptr, self.bidx, self.bw0, self.bbin = t(bptr.astype(np.int32)), t(bidx), t(bw0), t(bbin)
        self.schedT = t(np.argsort(-np.diff(WT.indptr), kind="stable").astype(np.int32))
        self.nnz = W0.nnz
        # the per-edge gain partials accumulate in the transposed product's chunk-padded order (its pads never read)
        realT = np.flatnonzero(bidx >= 0)
        pg, sg = _segments(bbin[realT].astype(np.int64), K * K)
        self.permG, self.segG = t(realT[pg].astype(np.int32)), t(sg)
        pt, st = _segments(ut, K)
        self.permT, self.segT = t(pt), t(st)
        self.unit_type = t(ut.astype(np.int32))
        ro = -np.ones(n, np.int32)
        ro[np.asarray(readout)] = np.arange(self.R, dtype=np.int32)
        self.readout = t(ro)
        self.P = torch.as_tensor(P, dtype=torch.float32, device=dev).contiguous()
        self.target = torch.as_tensor(target, dtype=torch.float32, device=dev).contiguous()
        z = lambda *s: torch.zeros(*s, device=dev)
        self.a, self.b = z(n), z(n)
        self.fw = z(len(fidx))
        self.bw = z(len(bidx))
        self.a_ad = float(math.exp(-c.dt / c.adapt_tau))
        self.f_fwd = (ct.c_float * 3)(c.gain_rr, c.adapt_gain, self.a_ad)
        self.loss_parts = z(self.T, self.R)
        self.loss_out = z(1)
        self.grid = grid or torch.cuda.get_device_properties(dev).multi_processor_count * 16
        nseg = (self.T + self.checkpoint - 1) // self.checkpoint
        self.ck_v = z(nseg, n, NL)
        self.ck_a = z(nseg, n, NL)
        C = self.checkpoint
        seg = lambda: (z(C + 1, n, NL), z(C + 1, n, NL), z(C + 1, n, NL), z(C, n, NL))
        # a segment's states and inputs: set 0 holds the forward's last segment; with several segments a second set
        # lets the backward of one segment overlap the recompute of the one before it (on a side stream)
        self._sets = [seg()] + ([seg()] if nseg > 1 else [])
        self.S_v, self.S_a, self.S_dr, self.S_inp = self._sets[0]
        self._side = torch.cuda.Stream(device=dev) if nseg > 1 else None
        self.lv, self.la, self.ly = z(n, NL), z(n, NL), z(n, NL)
        self.ly2 = z(n, NL)                      # ly of the next step, written by the backward launch
        self.ga, self.gb, self.eg = z(n), z(n), z(len(bidx))
        self._params = None
        self._last_segment_ready = False

    # ------------------------------------------------------------------ launches
    def _call(self, name, tensors, *args):
        arr = (ct.c_void_p * len(tensors))(*(None if x is None else x.data_ptr() for x in tensors))
        err = getattr(_lib(self.device), name)(arr, *args, torch.cuda.current_stream(self.device).cuda_stream)
        if err:
            raise RuntimeError(f"{name}: HIP error {err}")

    def _set_params(self, tau, bias, gain):
        tau = tau.detach().to(self.device, torch.float32).contiguous()
        bias = bias.detach().to(self.device, torch.float32).contiguous()
        gain = gain.detach().to(self.device, torch.float32).reshape(-1).contiguous()
        if tau.numel() != self.K or bias.numel() != self.K or gain.numel() != self.K * self.K:
            raise ValueError("tau and bias need one value per type, gain one per type pair")
        self._params = (tau, bias, gain)
        self._call("launch_adj_params", [tau, bias, self.unit_type, self.a, self.b], self.n, float(self.c.dt))
        self._call("launch_adj_weights", [gain, self.fbin, self.fw0, self.fw], self.fw.numel())
        self._call("launch_adj_weights", [gain, self.bbin, self.bw0, self.bw], self.bw.numel())

    def _step(self, t, v, ad, dr, v1, ad1, dr1, inp=None):
        self._call("launch_adj_forward",
                   [self.fptr, self.fidx, self.sched, self.fw, self.a, self.b, v, ad, dr, self.P[t], v1, ad1, dr1, inp,
                    self.readout, self.target[t], self.loss_parts[t]],
                   self.n, self.f_fwd, self.grid)

    # ------------------------------------------------------------------ forward
    def loss(self, tau, bias, gain) -> torch.Tensor:
        """Forward over T steps; returns the loss (a 1-element device tensor) and keeps the checkpoints."""
        self._set_params(tau, bias, gain)
        n, C = self.n, self.checkpoint
        v = [torch.zeros(n, NL, device=self.device) for _ in range(2)]
        ad = [torch.zeros(n, NL, device=self.device) for _ in range(2)]
        dr = [torch.empty(n, NL, device=self.device) for _ in range(2)]
        self._call("launch_adj_dr", [v[0], self.b, dr[0]], n)
        # the last segment's steps write straight into the segment buffers (with their inputs), so gradients() starts
        # without recomputing it; with checkpoint >= T that is the whole trajectory and nothing is recomputed
        t_last = ((self.T - 1) // C) * C
        for t in range(self.T):
            cur, nxt = t % 2, (t + 1) % 2
            if t % C == 0:
                self.ck_v[t // C].copy_(v[cur])
                self.ck_a[t // C].copy_(ad[cur])
            if t < t_last:
                self._step(t, v[cur], ad[cur], dr[cur], v[nxt], ad[nxt], dr[nxt])
                continue
            k = t - t_last
            if k == 0:
                self.S_v[0].copy_(v[cur])
                self.S_a[0].copy_(ad[cur])
                self._call("launch_adj_dr", [self.S_v[0], self.b, self.S_dr[0]], n)
            self._step(t, self.S_v[k], self.S_a[k], self.S_dr[k], self.S_v[k + 1], self.S_a[k + 1], self.S_dr[k + 1],
                       self.S_inp[k])
        self._last_segment_ready = True
        self._call("launch_loss_total", [self.loss_parts, self.loss_out], self.T * self.R,
                   1.0 / (self.T * NL * self.R))
        return self.loss_out

    def readout_trajectory(self, tau, bias, gain) -> torch.Tensor:
        """(T, R, 32) readout rates clip(v_t + b, 0, 1) for t = 1..T of a forward run (used to make teacher
        targets); the loss computed alongside is against the current targets and is discarded."""
        self._set_params(tau, bias, gain)
        n = self.n
        ro = torch.nonzero(self.readout >= 0).flatten()
        order = torch.argsort(self.readout[ro])
        ro = ro[order]
        v = [torch.zeros(n, NL, device=self.device) for _ in range(2)]
        ad = [torch.zeros(n, NL, device=self.device) for _ in range(2)]
        dr = [torch.empty(n, NL, device=self.device) for _ in range(2)]
        self._call("launch_adj_dr", [v[0], self.b, dr[0]], n)
        out = torch.empty(self.T, self.R, NL, device=self.device)
        for t in range(self.T):
            cur, nxt = t % 2, (t + 1) % 2
            self._step(t, v[cur], ad[cur], dr[cur], v[nxt], ad[nxt], dr[nxt])
            out[t] = (v[nxt][ro] + self.b[ro][:, None]).clamp(0.0, 1.0)
        return out

    # ------------------------------------------------------------------ backward
    def gradients(self):
        """(d loss/d tau, d loss/d bias, d loss/d gain) at the parameters of the last `loss` call.

        Segments are walked last to first.  Each one but the forward's last is recomputed from its checkpoint (the
        forward's launches: the same bits) into one of two segment buffers, on a side stream, while the backward of
        the segment after it runs on the current stream from the other buffer; events order a buffer's reuse after
        the backward that read it.  The backward's launches, and so its additions, are in the same order either way."""
        if self._params is None:
            raise RuntimeError("call loss() first")
        n, C, T = self.n, self.checkpoint, self.T
        scale = 2.0 / (T * NL * self.R)
        f_back = (ct.c_float * 4)(self.c.adapt_gain, self.a_ad, scale, self.c.gain_rr)
        for buf in (self.lv, self.la, self.ga, self.gb, self.eg):
            buf.zero_()
        nseg = (T + C - 1) // C
        main = torch.cuda.current_stream(self.device)
        if self._side is not None:
            self._side.wait_stream(main)         # the forward's checkpoints and parameters first
        freed = [None, None]                     # per buffer set: the event after the last backward that read it
        lys = [self.ly, self.ly2]                # ly of the current step, and of the next one
        ready = {}                               # per segment: the event after its recompute

        def recompute(s, which):
            t0, t1 = s * C, min(T, (s + 1) * C)
            S_v, S_a, S_dr, S_inp = self._sets[which]
            stream = self._side if self._side is not None else main
            with torch.cuda.stream(stream):
                if freed[which] is not None:
                    stream.wait_event(freed[which])
                S_v[0].copy_(self.ck_v[s])
                S_a[0].copy_(self.ck_a[s])
                self._call("launch_adj_dr", [S_v[0], self.b, S_dr[0]], n)
                for k in range(t1 - t0):
                    self._step(t0 + k, S_v[k], S_a[k], S_dr[k], S_v[k + 1], S_a[k + 1], S_dr[k + 1], S_inp[k])
                ev = torch.cuda.Event()
                ev.record(stream)
            ready[s] = ev

        last = nseg - 1
        if not self._last_segment_ready:
            recompute(last, 0)
        self._last_segment_ready = False
        for s in reversed(range(nseg)):
            which = (last - s) % 2
            if s > 0:
                recompute(s - 1, (last - s + 1) % 2)        # runs beside this segment's backward
            if s in ready:
                main.wait_event(ready[s])
            S_v, S_a, S_dr, S_inp = self._sets[which]
            t0, t1 = s * C, min(T, (s + 1) * C)
            if s == last:
                # the loss term of the last step, T: lambda_v_T = loss_scale (r_T - target_T) gate_T; it also
                # writes ly of step T - 1
                lys[0].zero_()
                self._back_rows(t1 - t0, T, f_back, None, S_v, S_dr, lys[0], lys[1])
                lys.reverse()
            for k in reversed(range(t1 - t0)):
                t = t0 + k
                # step t from ly_t: also step t's tau term and ly of step t - 1 (one launch per step)
                self._back_rows(k, t, f_back, self.eg, S_v, S_dr, lys[0], lys[1], S_inp[k])
                lys.reverse()
            ev = torch.cuda.Event()
            ev.record(main)
            freed[which] = ev
        tau, bias, gain = self._params
        K = self.K
        ga_t, gb_t, gg = (torch.empty(K, device=self.device), torch.empty(K, device=self.device),
                          torch.empty(K * K, device=self.device))
        self._call("launch_segment_sum", [self.ga, None, self.permT, self.segT, ga_t], K)
        self._call("launch_segment_sum", [self.gb, None, self.permT, self.segT, gb_t], K)
        self._call("launch_segment_sum", [self.eg, self.bw0, self.permG, self.segG, gg], K * K)
        a_k = torch.exp(-self.c.dt / tau)
        g_tau = ga_t * a_k * (self.c.dt / (tau * tau))
        return g_tau, gb_t, gg.view(K, K)

    def _back_rows(self, k, t, f_back, eg=None, S_v=None, S_dr=None, ly=None, ly_next=None, inp=None):
        """The transposed product and the rest of step t's adjoint (states of step t: row k of the segment buffers
        S_v, S_dr, by default the first set; ly of step t, by default self.ly); with eg, the same gathers also add
        each edge's sum over movies of ly[row] * dr_t[col] to eg (chunk-padded W^T order); with inp (step t's
        inputs) step t's tau term goes to ga, and with ly_next ly of step t - 1 is written there."""
        S_v = self.S_v if S_v is None else S_v
        S_dr = self.S_dr if S_dr is None else S_dr
        ly = self.ly if ly is None else ly
        target = self.target[t - 1] if t >= 1 else None
        self._call("launch_adj_back_rows",
                   [self.bptr, self.bidx, self.schedT, self.bw, self.a, self.b, ly, S_v[k], S_dr[k],
                    self.lv, self.la, self.readout, target, self.gb, eg, ly_next, inp, self.ga],
                   self.n, f_back, self.grid)


class OpticLossFn(torch.autograd.Function):
    """loss = OpticLossFn.apply(tau, bias, gain, problem): forward and the hand-written adjoint as one autograd node."""

    @staticmethod
    def forward(ctx, tau, bias, gain, problem: OpticAdjoint):
        ctx.problem = problem
        return problem.loss(tau, bias, gain).clone().reshape(())

    @staticmethod
    def backward(ctx, grad_out):
        g_tau, g_b, g_g = ctx.problem.gradients()
        return grad_out * g_tau, grad_out * g_b, grad_out * g_g, None
