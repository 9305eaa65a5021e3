"""The reference's CPU brain step and optic frame written with torch float32 CPU ops in the reference's order
(flyverse/brain.py:464-492, 679-692, 830-870; flyverse/optic.py:496-604; flyverse/fly.py:461-507) - TEST
INFRASTRUCTURE ONLY: the pin of the oracle's composite sequences (tests/test_composites.py) and the torch half of
bench.py's CPU baseline; the product never imports it.

The Poisson draw is injected (the caller passes the uniforms that replace torch.rand(generator)).  RefOptic's sparse
products are `mv`: chunked_mv, the chunked order every CSR product of this build defines (for bit-for-bit parity with
the oracle), or torch_mv, the reference's own torch sparse CSR product (optic.py:161-165; for timing the reference
computation).
"""
import math

import numpy as np
import scipy.sparse as sp
import torch


def chunked_mv(M: sp.csr_matrix, x: torch.Tensor) -> torch.Tensor:
    """(B, rows) = M @ x for x (B, cols) in float32: each row's stored entries in chunks of 16, a chunk folded from +0
    with every product rounded, the chunk sums folded from +0 (the order of every CSR product of this build)."""
    ptr, idx, w = M.indptr.astype(np.int64), M.indices.astype(np.int64), M.data.astype(np.float32)
    lens = np.diff(ptr)
    nch = (lens + 15) // 16
    B, rows = x.shape[0], M.shape[0]
    y = torch.zeros(B, rows)
    for c in range(int(nch.max()) if rows else 0):
        s = torch.zeros(B, rows)
        for j in range(16):
            k = c * 16 + j
            ok = torch.from_numpy(k < lens)
            e = np.where(k < lens, ptr[:-1] + k, 0)
            p = torch.from_numpy(w[e])[None] * x[:, torch.from_numpy(idx[e])]
            s = torch.where(ok[None], s + p, s)
        y = torch.where(torch.from_numpy(c < nch)[None], y + s, y)
    return y


def torch_mv(M: sp.csr_matrix):
    """The reference's `_mv` (flyverse/optic.py:161-165) for scipy CSR M: a function x (B, cols) -> (B, rows) through
    a torch sparse CSR tensor."""
    Mt = torch.sparse_csr_tensor(torch.from_numpy(M.indptr.astype(np.int64)), torch.from_numpy(M.indices.astype(np.int64)),
                                 torch.from_numpy(M.data.astype(np.float32)), size=M.shape)
    return lambda x: (Mt @ x.T.contiguous()).T


class RefBrain:
    """flyverse/brain.py's Brain state (:464-492) and CPU inference step (:830-870), event-driven input, torch ops."""

    def __init__(self, W: sp.csr_matrix, B: int, lp, active: np.ndarray):
        N = W.shape[0]
        Wc = sp.csc_matrix(W)
        Wc.sum_duplicates()
        Wc.sort_indices()
        self._out_ptr = torch.from_numpy(Wc.indptr.astype(np.int64))
        self._out_post = torch.from_numpy(Wc.indices.astype(np.int64))
        self._out_w = torch.from_numpy(Wc.data.astype(np.float32))
        self.p, self.n, self.B = lp, N, B
        z = lambda: torch.zeros(B, N)
        self.g, self.refrac, self.drive, self.poisson_p = z(), z(), z(), z()
        self.rate, self.spikes, self.spike_counts, self.adapt = z(), z(), z(), z()
        self.v = torch.full((B, N), lp.v_rest)
        self.res = torch.ones(B, N)
        self.active = torch.from_numpy(np.asarray(active, np.float32))
        self.n_delay = max(1, int(round(lp.delay / lp.dt)))
        self.spike_buf = torch.zeros(self.n_delay, B, N)
        self.buf_pos = 0
        self._a_m = math.exp(-lp.dt / lp.tau_m)
        self._a_s = math.exp(-lp.dt / lp.tau_syn)
        self._a_r = math.exp(-lp.dt / lp.rate_tau)
        self._a_ad = math.exp(-lp.dt / lp.adapt_tau) if lp.adapt_jump > 0 else 0.0
        self._std_on = lp.std_u > 0
        self._a_std = math.exp(-lp.dt / lp.std_tau) if self._std_on else 1.0
        self.std_u_vec = torch.full((N,), float(lp.std_u))

    def _add_synaptic_input(self, x):
        """brain.py:679-692, the CPU event-driven branch."""
        nz = torch.nonzero(x)
        if nz.shape[0] == 0:
            return
        bidx, pre = nz[:, 0], nz[:, 1]
        start = self._out_ptr[pre]
        cnt = self._out_ptr[pre + 1] - start
        total = int(cnt.sum())
        if total == 0:
            return
        seg = torch.repeat_interleave(torch.arange(nz.shape[0]), cnt, output_size=total)
        first = torch.repeat_interleave(torch.cumsum(cnt, 0) - cnt, cnt, output_size=total)
        e = start[seg] + torch.arange(total) - first
        tgt = self._out_post[e] + bidx[seg] * self.n
        self.g.view(-1).index_add_(0, tgt, self._out_w[e] * x[bidx[seg], pre[seg]])

    def step(self, rnd=None):
        """One step of brain.py:830-868; rnd replaces torch.rand(generator) when Poisson forcing is on."""
        p = self.p
        self.g.mul_(self._a_s)
        self._add_synaptic_input(self.spike_buf[self.buf_pos])
        target = p.v_rest + self.g + self.drive - self.adapt
        torch.add(target, (self.v - target) * self._a_m, out=self.v)
        in_ref = self.refrac > 0
        torch.where(in_ref, torch.full_like(self.v, p.v_reset), self.v, out=self.v)
        torch.clamp(self.refrac - p.dt, min=0.0, out=self.refrac)
        torch.mul((self.v >= p.v_th).float(), self.active, out=self.spikes)
        spikes = self.spikes
        if rnd is not None:
            forced = (rnd < self.poisson_p).float()
            torch.maximum(spikes, forced, out=spikes)
        fired = spikes > 0
        torch.where(fired, torch.full_like(self.v, p.v_reset), self.v, out=self.v)
        torch.where(fired, torch.full_like(self.refrac, p.t_ref), self.refrac, out=self.refrac)
        if p.adapt_jump > 0:
            self.adapt.mul_(self._a_ad).add_(spikes, alpha=p.adapt_jump)
        if self._std_on:
            self.spike_buf[self.buf_pos] = spikes * self.res
            torch.where(fired, self.res * (1 - self.std_u_vec), self.res, out=self.res)
            torch.sub(1.0, (1.0 - self.res) * self._a_std, out=self.res)
        else:
            self.spike_buf[self.buf_pos] = spikes
        self.buf_pos = (self.buf_pos + 1) % self.n_delay
        self.spike_counts.add_(spikes)
        self.rate.mul_(self._a_r).add_(spikes * ((1 - self._a_r) * 1000.0 / p.dt))


class RefClockedBrain(RefBrain):
    """Brain._step_clocked on the torch branch (flyverse/brain.py:618-664, 898-957): per-neuron clocks with the phase
    coefficients of set_clocks, one accumulator per clock longer than a step, and each clock's product W_k x formed
    into zeros with index_add_ in ascending presynaptic order (the product order this build defines) before g.add_."""

    def set_clocks(self, k, W):
        p = self.p
        k = np.asarray(k, np.int64)
        self.K = int(np.lcm.reduce(np.unique(k)))
        kt = torch.from_numpy(k.astype(np.float32))
        self._phase = []
        for ph in range(self.K):
            u = torch.from_numpy(((ph % k) == 0).astype(np.float32))
            one = torch.ones_like(u)
            coef = lambda tau: torch.where(u > 0, torch.exp(-p.dt * kt / tau), one)
            self._phase.append({"u": u, "a_m": coef(p.tau_m), "a_s": coef(p.tau_syn), "a_r": coef(p.rate_tau),
                                "a_ad": coef(p.adapt_tau) if p.adapt_jump > 0 else torch.zeros_like(u),
                                "a_std": coef(p.std_tau) if self._std_on else one, "dt": u * kt * p.dt,
                                "rate_gain": u * (1 - torch.exp(-p.dt * kt / p.rate_tau)) * 1000.0 / (kt * p.dt),
                                "pois": u * kt})
        self._W_k, self._acc = {}, {}
        for kk in np.unique(k):
            Wk = sp.csc_matrix(sp.diags((k == kk).astype(np.float32)) @ sp.csr_matrix(W))
            Wk.eliminate_zeros()
            Wk.sort_indices()
            self._W_k[int(kk)] = (torch.from_numpy(Wk.indptr.astype(np.int64)), torch.from_numpy(Wk.indices.astype(np.int64)),
                                  torch.from_numpy(Wk.data.astype(np.float32)))
            if kk > 1:
                self._acc[int(kk)] = torch.zeros(self.B, self.n)
        self.step_count = 0

    def _product(self, Wk, x):
        ptr, post, w = Wk
        y = torch.zeros(self.B * self.n)
        nz = torch.nonzero(x)
        if nz.shape[0]:
            bidx, pre = nz[:, 0], nz[:, 1]
            start = ptr[pre]
            cnt = ptr[pre + 1] - start
            total = int(cnt.sum())
            if total:
                seg = torch.repeat_interleave(torch.arange(nz.shape[0]), cnt, output_size=total)
                first = torch.repeat_interleave(torch.cumsum(cnt, 0) - cnt, cnt, output_size=total)
                e = start[seg] + torch.arange(total) - first
                y.index_add_(0, post[e] + bidx[seg] * self.n, w[e] * x[bidx[seg], pre[seg]])
        return y.view(self.B, self.n)

    def step(self, rnd=None):
        p = self.p
        c = self._phase[self.step_count % self.K]
        x = self.spike_buf[self.buf_pos]
        self.g.mul_(c["a_s"])
        for kk, Wk in self._W_k.items():
            if kk == 1:
                self.g.add_(self._product(Wk, x))
            else:
                acc = self._acc[kk]
                acc.add_(x)
                if (self.step_count % kk) == 0:
                    self.g.add_(self._product(Wk, acc))
                    acc.zero_()
        target = p.v_rest + self.g + self.drive - self.adapt
        torch.add(target, (self.v - target) * c["a_m"], out=self.v)
        in_ref = self.refrac > 0
        torch.where(in_ref, torch.full_like(self.v, p.v_reset), self.v, out=self.v)
        torch.clamp(self.refrac - c["dt"], min=0.0, out=self.refrac)
        torch.mul((self.v >= p.v_th).float() * c["u"], self.active, out=self.spikes)
        spikes = self.spikes
        if rnd is not None:
            forced = (rnd < self.poisson_p * c["pois"]).float()
            torch.maximum(spikes, forced, out=spikes)
        fired = spikes > 0
        torch.where(fired, torch.full_like(self.v, p.v_reset), self.v, out=self.v)
        torch.where(fired, torch.full_like(self.refrac, p.t_ref), self.refrac, out=self.refrac)
        if p.adapt_jump > 0:
            self.adapt.mul_(c["a_ad"]).add_(spikes, alpha=p.adapt_jump)
        if self._std_on:
            self.spike_buf[self.buf_pos] = spikes * self.res
            torch.where(fired, self.res * (1 - self.std_u_vec), self.res, out=self.res)
            torch.sub(1.0, (1.0 - self.res) * c["a_std"], out=self.res)
        else:
            self.spike_buf[self.buf_pos] = spikes
        self.buf_pos = (self.buf_pos + 1) % self.n_delay
        self.spike_counts.add_(spikes)
        self.rate.mul_(c["a_r"]).add_(spikes * c["rate_gain"])
        self.step_count += 1


class RefOptic:
    """flyverse/optic.py's OpticLobe frame on the CPU branch (:496-604), torch ops, products in the chunked order."""

    def __init__(self, m: dict, op, retina, B: int, n: int, products: str = "chunked"):
        self.m, self.p, self.B, self.n = m, op, B, n
        if products == "chunked":
            self._mv = {k: (lambda x, M=m[k]: chunked_mv(M, x)) for k in ("M_rr", "M_rp", "M_rs", "M_sr")}
        elif products == "torch":
            self._mv = {k: torch_mv(m[k]) for k in ("M_rr", "M_rp", "M_rs", "M_sr")}
        else:
            raise ValueError("products must be chunked or torch")
        n_rate = len(m["rate_idx"])
        self.v = torch.zeros(B, n_rate)
        self.adapt = torch.zeros(B, n_rate)
        self.b_vec = torch.from_numpy(m["b"])
        self._a = torch.from_numpy(m["a"])
        self._a_ad = m["a_ad"]
        self.r0 = self.b_vec[None].clone()                    # relax(), optic.py:553-556
        n_col = retina.n_col
        pr_col, fam = retina.pr_column, retina.pr_family
        cnt = np.zeros((n_col, 5), np.float32)
        np.add.at(cnt, (pr_col, fam), 1.0)
        rows = pr_col * 5 + fam
        # the averaging matrix and its masks as optic.py:315-325 builds them
        self.avg = sp.csr_matrix((1.0 / cnt[pr_col, fam], (rows, np.arange(len(pr_col)))), shape=(n_col * 5,
                                                                                                   len(pr_col)))
        self.avg.sort_indices()
        self._mv["avg"] = (lambda x, M=self.avg: chunked_mv(M, x)) if products == "chunked" else torch_mv(self.avg)
        self.has_t = torch.from_numpy((cnt > 0).reshape(-1))
        self.pr_cell_t = torch.from_numpy(rows)
        self.pr_column_t = torch.from_numpy(pr_col)
        self.sens = torch.from_numpy(retina.pr_sens)
        self.I_lp = torch.zeros(B, n_col * 5)
        self.I_mean = torch.zeros(B, n_col * 5)
        self.contrast = torch.zeros(B, n_col * 5)
        self._fresh = torch.ones(B, dtype=torch.bool)
        self.delta_rate = torch.zeros(B, n_rate)
        self.spk_idx_t = torch.from_numpy(m["spk_idx"])
        self._pending_ms = 0.0

    def photoreceptor_activity(self, rad, dt_ms):
        p = self.p
        I_pr = (rad[:, self.pr_column_t, :] * self.sens[None]).sum(-1)
        I = self._mv["avg"](I_pr)
        fresh = self._fresh[:, None]
        torch.where(fresh, I, self.I_lp, out=self.I_lp)
        torch.where(fresh, I, self.I_mean, out=self.I_mean)
        self._fresh[:] = False
        a_lp = float(np.exp(-dt_ms / p.tau_lp_ms))
        a_ad = float(np.exp(-dt_ms / p.tau_adapt_ms))
        torch.add(a_lp * self.I_lp, (1 - a_lp) * I, out=self.I_lp)
        torch.add(a_ad * self.I_mean, (1 - a_ad) * self.I_lp, out=self.I_mean)
        contrast = ((self.I_lp - self.I_mean) / (self.I_mean + p.eps)).clamp(-1.0, p.contrast_clip)
        torch.where(self.has_t[None], contrast, torch.zeros_like(contrast), out=self.contrast)
        return self.contrast[:, self.pr_cell_t]

    def rates(self):
        return (self.v + self.b_vec[None]).clamp(0.0, 1.0)

    def _substep(self, pr_input, spk_input):
        p = self.p
        dr = self.rates() - self.b_vec[None]
        inp = p.gain_rr * self._mv["M_rr"](dr) + pr_input - p.adapt_gain * self.adapt
        inp = inp + spk_input
        torch.add(inp, (self.v - inp) * self._a[None], out=self.v)
        torch.add(dr, (self.adapt - dr) * self._a_ad, out=self.adapt)

    def step_frame(self, rad, spk_rate_hz, frame_ms):
        a_pr = self.photoreceptor_activity(rad, frame_ms)
        s = (spk_rate_hz[:, self.spk_idx_t] / 100.0).clamp(0, 3)
        total = self._pending_ms + frame_ms
        steps = int(np.floor((total + 1e-9) / self.p.dt_ms))
        self._pending_ms = max(0.0, total - steps * self.p.dt_ms)
        if steps:
            pr_input = self.p.gain_in * self._mv["M_rp"](a_pr)
            spk_input = self.p.gain_fb * self._mv["M_rs"](s)
            for _ in range(steps):
                self._substep(pr_input, spk_input)
        torch.sub(self.rates(), self.r0, out=self.delta_rate)
        drive = torch.zeros(self.B, self.n)
        drive[:, self.spk_idx_t] = (self.p.gain_out_mv * self._mv["M_sr"](self.delta_rate)).clamp(
            -self.p.drive_clip_mv, self.p.drive_clip_mv)
        return drive


def fly_frame_steps(pending, ms, dt):
    """FlyBrain._step_plain's split (flyverse/fly.py:461-464): whole LIF steps and the carried remainder."""
    total = pending + ms
    steps = int(math.floor((total + 1e-9) / dt))
    return steps, max(0.0, total - steps * dt)
