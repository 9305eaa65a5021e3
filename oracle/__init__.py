"""CPU oracle for the brain step — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package, and only as the checker
(or, in bench.py, as the timed CPU baseline); the product path (the HIP library and the flyhip package) never does.

It wraps oracle/oracle.c (a plain-C restatement of the reference arithmetic, built by oracle/Makefile) and composes
the reference's step sequences from it:
  * `brain_step`   — Brain._step_inference with the event-driven CPU synaptic input (flyverse/brain.py:818-870,
                      679-692), the Poisson draw replaced by the Philox uniforms of `uniform`;
  * `optic_frame`  — OpticLobe._step_frame_inference (flyverse/optic.py:563-604) with the substep of :525-551;
  * `fly_run`      — the FlyBrain frame loop (flyverse/fly.py:477-507): optic frame first, its drive copied into
                      the LIF, then the LIF steps of the frame.

Pinning (what checks each function against something other than itself), by function:
  philox4x32_10  Random123 known-answer vectors (tests/test_oracle.py).
  lif            closed forms of the exponential-Euler membrane, the fp32 refractory chain 2.2f - k*0.1f, the
                 threshold equality case, frozen/forced/STD/adaptation branches (tests/test_oracle.py), and a torch
                 float32 restatement of flyverse/brain.py:838-868 run op by op (tests/test_oracle.py).
  events         brute-force dense float64 x @ W.T within 1e-6 relative; bit-exact against an explicit
                 Python loop over nonzero(x) in (b, pre) order on small graphs (tests/test_oracle.py).
  spmv           scipy float64 M @ x within 1e-5 relative; the chunk order against a pure-Python fold
                 (tests/test_oracle.py); half rounding against numpy float16 casting.
  optic_dr/update  the torch float32 expressions of flyverse/optic.py:528-551 evaluated op by op (exact).
  photoreceptor  the torch float32 expressions of flyverse/optic.py:503-517 evaluated op by op on the reference's
                 averaging matrix (optic.py:320-325), its product folded in the CSR order (exact) and torch's own
                 sparse product within 1e-6; float64 numpy with that matrix within 1e-5 relative (tests/test_oracle.py).
  group_means    numpy float64 means.
"""
from __future__ import annotations

import ctypes as ct
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_I, _F, _U64, _U32, _L = ct.c_int, ct.c_float, ct.c_uint64, ct.c_uint32, ct.c_int64


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ct.CDLL(str(LIB_PATH))
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_set_threads.argtypes = [_I]
        L.orc_uniform1.argtypes = [_U64, _U32, _U32, _U32, _U32]
        L.orc_uniform1.restype = ct.c_float
        L.orc_uniform.argtypes = [_f32p, _I, _I, _U64, _U32, _U32]
        L.orc_uniform_at.argtypes = [_f32p, _i32p, _I, _U64, _U32, _U32, _U32]
        L.orc_lif.argtypes = [_I, _I, _I] + [_f32p] * 16
        L.orc_scale.argtypes = [_f32p, _L, _F]
        L.orc_events.argtypes = [_I, _I, _i64p, _i32p, _f32p, _f32p, _f32p]
        L.orc_spmv.argtypes = [_I, _I, _I, _i64p, _i32p, _f32p, _f32p, _f32p, _I]
        L.orc_optic_dr.argtypes = [_I, _I, _f32p, _f32p, _f32p]
        L.orc_optic_update.argtypes = [_I, _I] + [_f32p] * 8 + [_F, _F, _F]
        L.orc_photoreceptor.argtypes = [_I, _I, _I, _f32p, _f32p, _i32p, _i32p, _i32p, _f32p, _f32p, _f32p, _u8p,
                                        _f32p, _f32p, _f32p]
        L.orc_group_means.argtypes = [_I, _I, _I, _I, _i32p, _i32p, _f32p, _f64p]
        L.orc_drive_out.argtypes = [_I, _I, _I, _f32p, _i32p, _F, _F, _f32p]
        L.orc_feedback.argtypes = [_I, _I, _I, _f32p, _i32p, _f32p]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ------------------------------------------------------------------ primitives
def set_threads(n: int) -> None:
    """Threads of the oracle's parallel loops (each writes disjoint elements: the results do not change)."""
    lib().orc_set_threads(int(n))


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def uniform(batch, n, seed, step, stream=0):
    out = np.empty((batch, n), np.float32)
    lib().orc_uniform(out, batch, n, int(seed) & (2**64 - 1), int(step) & 0xFFFFFFFF, stream)
    return out


def uniform_at(idx, seed, step, fly=0, stream=0):
    """The uniforms of cells idx only (the same values uniform() gives at those cells)."""
    idx = np.ascontiguousarray(idx, np.int32)
    out = np.empty(len(idx), np.float32)
    lib().orc_uniform_at(out, idx, len(idx), int(seed) & (2**64 - 1), int(step) & 0xFFFFFFFF, fly, stream)
    return out


def uniform1(seed, step, neuron, fly=0, stream=0) -> float:
    return float(lib().orc_uniform1(int(seed) & (2**64 - 1), step & 0xFFFFFFFF, neuron, fly, stream))


def lif(state: dict, flags: int, P, rnd=None, phase=None):
    """In-place fused LIF update on a dict of float32 arrays (keys as Brain's tensors; `out` = the spike_buf slot)."""
    s = state
    B, N = s["v"].shape
    P = _f32(P)
    rnd = s["poisson_p"] if rnd is None else _f32(rnd)
    phase = P if phase is None else _f32(phase)
    lib().orc_lif(N, B, flags, s["v"], s["g"], s["drive"], s["adapt"], s["refrac"], s["active"], s["poisson_p"], rnd,
                  s["res"], s["std_u"], s["spikes"], s["out"], s["rate"], s["spike_counts"], P, phase)


def scale(x, a):
    lib().orc_scale(x.reshape(-1), x.size, np.float32(a))


def events(csc_ptr, csc_post, csc_w, x, g):
    B, N = x.shape
    lib().orc_events(N, B, np.ascontiguousarray(csc_ptr, np.int64), np.ascontiguousarray(csc_post, np.int32),
                     _f32(csc_w), _f32(x), g)


def spmv(ptr, idx, w, x, half=False):
    rows = len(ptr) - 1
    B, cols = x.shape
    y = np.empty((B, rows), np.float32)
    lib().orc_spmv(rows, cols, B, np.ascontiguousarray(ptr, np.int64), np.ascontiguousarray(idx, np.int32), _f32(w),
                   _f32(x), y, int(half))
    return y


def optic_dr(v, b):
    B, n = v.shape
    dr = np.empty_like(v)
    lib().orc_optic_dr(n, B, _f32(v), _f32(b), dr)
    return dr


def optic_update(v, adapt, dr, y, pr, spk, a, b, gain_rr, adapt_gain, a_ad):
    B, n = v.shape
    lib().orc_optic_update(n, B, v, adapt, dr, _f32(y), _f32(pr), _f32(spk), _f32(a), _f32(b), np.float32(gain_rr),
                           np.float32(adapt_gain), np.float32(a_ad))


def identity_cells(n_pr):
    """The cell tables of a retina with one photoreceptor per column, all of family 0 (cell = column * 5)."""
    return {"cell_ptr": np.repeat(np.arange(n_pr + 1, dtype=np.int32), [1] + [5] * n_pr)[:5 * n_pr + 1],
            "cell_pr": np.arange(n_pr, dtype=np.int32), "cell_w": np.ones(n_pr, np.float32),
            "pr_column": np.arange(n_pr, dtype=np.int32)}


def photoreceptor(rad, sens, I_lp, I_mean, fresh, coefs, cells=None, contrast=None):
    """rad (B, n_col, 4); I_lp, I_mean (B, n_col*5) updated in place; returns a_pr (B, n_pr).  `cells` holds the
    averaging tables (cell_ptr, cell_pr, cell_w, pr_column), by default those of identity_cells."""
    B, n_col = rad.shape[:2]
    n_pr = sens.shape[0]
    cells = identity_cells(n_pr) if cells is None else cells
    if I_lp.shape != (B, 5 * n_col) or len(cells["cell_ptr"]) != 5 * n_col + 1:
        raise ValueError("photoreceptor state must be (batch, 5 * columns)")
    a_pr = np.zeros((B, n_pr), np.float32)
    contrast = np.empty((B, 5 * n_col), np.float32) if contrast is None else contrast
    lib().orc_photoreceptor(n_col, n_pr, B, _f32(rad), _f32(sens), np.ascontiguousarray(cells["pr_column"], np.int32),
                            np.ascontiguousarray(cells["cell_ptr"], np.int32),
                            np.ascontiguousarray(cells["cell_pr"], np.int32), _f32(cells["cell_w"]), I_lp, I_mean,
                            fresh, _f32(coefs), contrast, a_pr)
    return a_pr


def group_means(ptr, idx, rates, stride=None):
    B, n = rates.shape
    groups = len(ptr) - 1
    stride = groups if stride is None else stride
    out = np.zeros((B, stride), np.float64)
    lib().orc_group_means(n, groups, B, stride, np.ascontiguousarray(ptr, np.int32), np.ascontiguousarray(idx, np.int32),
                          _f32(rates), out)
    return out


def drive_out(y_sr, spk_idx, n, gain_out, clip):
    B = y_sr.shape[0]
    drive = np.empty((B, n), np.float32)
    lib().orc_drive_out(n, len(spk_idx), B, _f32(y_sr), np.ascontiguousarray(spk_idx, np.int32), np.float32(gain_out),
                        np.float32(clip), drive)
    return drive


def feedback(rate, spk_idx):
    B, n = rate.shape
    s = np.empty((B, len(spk_idx)), np.float32)
    lib().orc_feedback(n, len(spk_idx), B, _f32(rate), np.ascontiguousarray(spk_idx, np.int32), s)
    return s


# ------------------------------------------------------------------ composite reference sequences
def lif_params_vector(p) -> np.ndarray:
    """[v_rest, v_reset, v_th, a_m, dt, t_ref, adapt_jump, a_ad, a_std, a_r, rate_gain] (flyverse/brain.py:493-520),
    coefficients in host double, then float32."""
    import math
    a_m = math.exp(-p.dt / p.tau_m)
    a_r = math.exp(-p.dt / p.rate_tau)
    a_ad = math.exp(-p.dt / p.adapt_tau) if p.adapt_jump > 0 else 0.0
    a_std = math.exp(-p.dt / p.std_tau) if p.std_u > 0 else 1.0
    return np.array([p.v_rest, p.v_reset, p.v_th, a_m, p.dt, p.t_ref, p.adapt_jump, a_ad, a_std, a_r,
                     (1 - a_r) * 1000.0 / p.dt], dtype=np.float32)


class BrainState:
    """The (B, N) state of flyverse/brain.py:464-492, as float32 numpy arrays, plus the spike delay ring."""
    FIELDS = ("v", "g", "refrac", "drive", "poisson_p", "rate", "spikes", "spike_counts", "adapt", "res")

    def __init__(self, B, N, n_delay, v_rest):
        self.v = np.full((B, N), v_rest, np.float32)
        for k in ("g", "refrac", "drive", "poisson_p", "rate", "spikes", "spike_counts", "adapt"):
            setattr(self, k, np.zeros((B, N), np.float32))
        self.res = np.ones((B, N), np.float32)
        self.active = np.ones(N, np.float32)
        self.std_u = np.zeros(N, np.float32)
        self.spike_buf = np.zeros((n_delay, B, N), np.float32)
        self.buf_pos = 0
        self.step_count = 0
        self.n_delay = n_delay


def brain_step(st: BrainState, csc, P, a_s, seed, flags, poisson_on=True, record=True):
    """One Brain._step_inference step on the CPU (flyverse/brain.py:830-870): g *= a_s; the event-driven input of
    the spikes emitted n_delay steps ago; the Poisson uniforms of step `step_count` (Philox, fly index = row); the
    fused update; the ring advances.  csc = (ptr int64, post int32, w float32), pre-major."""
    scale(st.g, a_s)
    events(csc[0], csc[1], csc[2], st.spike_buf[st.buf_pos], st.g)
    B, N = st.v.shape
    fl = flags | (1 if poisson_on else 0) | (4 if record else 0)
    rnd = uniform(B, N, seed, st.step_count) if poisson_on else st.poisson_p
    slot = st.spike_buf[st.buf_pos]
    d = dict(v=st.v, g=st.g, drive=st.drive, adapt=st.adapt, refrac=st.refrac, active=st.active,
             poisson_p=st.poisson_p, res=st.res, std_u=st.std_u, spikes=st.spikes, out=slot, rate=st.rate,
             spike_counts=st.spike_counts)
    lif(d, fl, P, rnd)
    st.buf_pos = (st.buf_pos + 1) % st.n_delay
    st.step_count += 1


class OpticState:
    def __init__(self, B, n_rate, n_cells):
        self.v = np.zeros((B, n_rate), np.float32)
        self.adapt = np.zeros((B, n_rate), np.float32)
        self.dr = np.zeros((B, n_rate), np.float32)
        self.I_lp = np.zeros((B, n_cells), np.float32)
        self.I_mean = np.zeros((B, n_cells), np.float32)
        self.fresh = np.ones(B, np.uint8)
        self.pending_ms = 0.0


def optic_frame(os_: OpticState, g, rad, rate, frame_ms):
    """OpticLobe._step_frame_inference (flyverse/optic.py:563-604) on graph `g` (a flyhip.synth.FlyGraph): returns the
    (B, N) drive written into the LIF (fly.py:492)."""
    a_pr = photoreceptor(rad, g.pr_sens, os_.I_lp, os_.I_mean, os_.fresh, g.pr_coefs(frame_ms), g.cells)
    s = feedback(rate, g.spk_idx)
    total = os_.pending_ms + frame_ms
    steps = int(np.floor((total + 1e-9) / g.optic.dt_ms))
    os_.pending_ms = max(0.0, total - steps * g.optic.dt_ms)
    B = rate.shape[0]
    if steps:
        pr_in = spmv(*g.W_rp, a_pr)
        pr_in = (np.float32(g.optic.gain_in) * pr_in).astype(np.float32)
        spk_in = spmv(*g.W_rs, s)
        spk_in = (np.float32(g.optic.gain_fb) * spk_in).astype(np.float32)
        dr = optic_dr(os_.v, g.b_vec)
        for _ in range(steps):
            y = spmv(*g.W_rr, dr)
            optic_update(os_.v, os_.adapt, dr, y, pr_in, spk_in, g.a_vec, g.b_vec, g.optic.gain_rr,
                         g.optic.adapt_gain, g.a_ad)
        os_.dr = dr
    dr = optic_dr(os_.v, g.b_vec)               # delta_rate = rates() - r0 with r0 = b (optic.py:599)
    y_sr = spmv(*g.W_sr, dr)
    return drive_out(y_sr, g.spk_idx, g.n, g.optic.gain_out_mv, g.optic.drive_clip_mv)


# ------------------------------------------------------------------ the whole fly (flyverse/fly.py:448-507)
class OptSpec:
    """Host-built optic-lobe inputs of one graph (the arrays flyhip.optic.optic_matrices produces)."""
    def __init__(self, m: dict, params, sens, n, cells=None):
        to = lambda M: (M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data.astype(np.float32))
        self.W_rr, self.W_rp, self.W_rs, self.W_sr = (to(m[k]) for k in ("M_rr", "M_rp", "M_rs", "M_sr"))
        self.a_vec, self.b_vec, self.a_ad = m["a"], m["b"], m["a_ad"]
        self.spk_idx = m["spk_idx"].astype(np.int32)
        self.pr_sens = np.ascontiguousarray(sens, np.float32)
        self.cells = identity_cells(len(self.pr_sens)) if cells is None else cells
        self.n_cells = len(self.cells["cell_ptr"]) - 1
        self.n_col = self.n_cells // 5
        self.optic = params
        self.n = n

    def pr_coefs(self, frame_ms):
        a_lp = float(np.exp(-frame_ms / self.optic.tau_lp_ms))
        a_ad = float(np.exp(-frame_ms / self.optic.tau_adapt_ms))
        return np.array([a_lp, 1 - a_lp, a_ad, 1 - a_ad, self.optic.eps, self.optic.contrast_clip], np.float32)


def _per_fly_uniform(B, n, seeds, step, stream):
    """seeds an int: one stream, fly index = row; a sequence: row b uses seeds[b] with fly index 0."""
    if np.isscalar(seeds):
        return uniform(B, n, int(seeds), step, stream)
    return np.stack([uniform(1, n, int(s), step, stream)[0] for s in seeds])


class FlyModel:
    """B flies on the CPU: LIF (event-driven reference order) plus, optionally, the optic lobe coupled per frame."""
    def __init__(self, n, B, P, a_s, n_delay, csc, active, seeds, optic: OptSpec | None = None, movie_seeds=None,
                 v_rest=-52.0, dt=0.1):
        self.st = BrainState(B, n, n_delay, v_rest)
        self.st.active = np.ascontiguousarray(active, np.float32)
        self.P, self.a_s, self.csc, self.seeds, self.dt = _f32(P), float(a_s), csc, seeds, dt
        self.optic = optic
        self.os = OpticState(B, len(optic.b_vec), optic.n_cells) if optic is not None else None
        self.movie_seeds = movie_seeds
        self.frame_index = 0
        self.B, self.n = B, n

    flags = 32                                    # frozen-at-rest skip, as Brain(cuda_compact=True)

    poisson_cells = None                          # set: draw uniforms only at these cells (the others get 1.0)
    clocks = None                                 # set by set_clocks

    def set_clocks(self, k, W, p):
        """Per-neuron clocks (flyverse/brain.py:618-664): k (N,) multipliers, W the (post, pre) weights, p the LIF
        parameters.  The phase coefficients restate brain.py:630-645 with torch float32 CPU ops; each clock's rows of
        W as a pre-major CSC (its products fold in ascending presynaptic order from +0); one accumulator per clock
        longer than one step."""
        import torch
        import scipy.sparse as sp
        k = np.asarray(k, np.int64)
        K = int(np.lcm.reduce(np.unique(k)))
        kt = torch.from_numpy(k.astype(np.float32))
        std_on = p.std_u > 0
        phases = []
        for ph in range(K):
            u = torch.from_numpy(((ph % k) == 0).astype(np.float32))
            one = torch.ones_like(u)
            coef = lambda tau: torch.where(u > 0, torch.exp(-p.dt * kt / tau), one)
            rows = [coef(p.tau_m), u * kt * p.dt, coef(p.adapt_tau) if p.adapt_jump > 0 else torch.zeros_like(u),
                    coef(p.std_tau) if std_on else one, coef(p.rate_tau),
                    u * (1 - torch.exp(-p.dt * kt / p.rate_tau)) * 1000.0 / (kt * p.dt), u, u * kt]
            phases.append((coef(p.tau_syn).numpy(), torch.stack(rows).numpy().astype(np.float32)))
        cscs = {}
        for kk in np.unique(k):
            Wk = (sp.diags((k == kk).astype(np.float32)) @ sp.csr_matrix(W)).tocsc()
            Wk.eliminate_zeros()
            Wk.sort_indices()
            cscs[int(kk)] = (Wk.indptr.astype(np.int64), Wk.indices.astype(np.int32), Wk.data.astype(np.float32))
        B, N = self.st.v.shape
        self.clocks = {"K": K, "phases": phases, "cscs": cscs,
                       "acc": {int(kk): np.zeros((B, N), np.float32) for kk in np.unique(k) if kk > 1}}

    def _clocked_input(self, x):
        """The synaptic input of Brain._step_clocked (flyverse/brain.py:903-921): g *= a_s of the phase; for each clock
        in ascending order, clock 1 adds W_1 x, a longer clock adds x to its accumulator and, on the steps its period
        divides, adds W_k acc (then clears it); each product y from +0, then g + y."""
        st, ck = self.st, self.clocks
        a_s, lifphase = ck["phases"][st.step_count % ck["K"]]
        st.g[:] = (st.g * a_s[None]).astype(np.float32)
        for kk, csc in ck["cscs"].items():
            if kk == 1:
                xin = x
            else:
                acc = ck["acc"][kk]
                acc += x
                if st.step_count % kk:
                    continue
                xin = acc
            y = np.zeros_like(st.g)
            events(csc[0], csc[1], csc[2], xin, y)
            st.g[:] = (st.g + y).astype(np.float32)
            if kk > 1:
                ck["acc"][kk][:] = 0.0
        return lifphase

    def _poisson_uniforms(self, B, N, step):
        if self.poisson_cells is None:
            return _per_fly_uniform(B, N, self.seeds, step, 0)
        rnd = np.ones((B, N), np.float32)              # never below a probability <= 1: those cells cannot be forced
        for b in range(B):
            seed, fly = (int(self.seeds), b) if np.isscalar(self.seeds) else (int(self.seeds[b]), 0)
            rnd[b, self.poisson_cells] = uniform_at(self.poisson_cells, seed, step, fly)
        return rnd

    def lif_steps(self, steps, poisson_on=None, record=True):
        st = self.st
        on = bool((st.poisson_p > 0).any()) if poisson_on is None else poisson_on
        B, N = st.v.shape
        if on and self.poisson_cells is not None and (st.poisson_p[:, np.setdiff1d(np.arange(N), self.poisson_cells)] > 0).any():
            raise ValueError("poisson_cells must hold every cell with a Poisson probability")
        for _ in range(steps):
            phase = None
            if self.clocks is None:
                scale(st.g, self.a_s)
                events(*self.csc, st.spike_buf[st.buf_pos], st.g)
            else:
                phase = self._clocked_input(st.spike_buf[st.buf_pos])
            rnd = self._poisson_uniforms(B, N, st.step_count) if on else st.poisson_p
            slot = st.spike_buf[st.buf_pos]
            d = dict(v=st.v, g=st.g, drive=st.drive, adapt=st.adapt, refrac=st.refrac, active=st.active,
                     poisson_p=st.poisson_p, res=st.res, std_u=st.std_u, spikes=st.spikes, out=slot, rate=st.rate,
                     spike_counts=st.spike_counts)
            flags = self.flags | (1 if on else 0) | (4 if record else 0) | (16 if phase is not None else 0)
            lif(d, flags, self.P, rnd, phase)
            st.buf_pos = (st.buf_pos + 1) % st.n_delay
            st.step_count += 1
            yield st

    def frame(self, steps, raster=None):
        """One frame: optic first (if any) with a fresh random movie, then `steps` LIF steps; appends spikes."""
        if self.optic is not None:
            ncol = self.optic.n_col
            rad = _per_fly_uniform(self.B, ncol * 4, self.movie_seeds, self.frame_index, 1).reshape(self.B, ncol, 4)
            self.st.drive[:] = optic_frame(self.os, self.optic, rad, self.st.rate, steps * self.dt)
        for st in self.lif_steps(steps):
            if raster is not None:
                raster.append(st.spikes.copy())
        self.frame_index += 1
