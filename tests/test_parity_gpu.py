"""GPU parity of every plain-path kernel against the CPU oracle, through the C ABI (flyhip.hip wraps libflyhip.so).

The bar: integer, index and spike work bit-exact; the LIF, pull, optic and CSR kernels bit-exact too, because the
oracle restates the same fp32 operation order (the reference CPU order for the LIF and the synaptic input, the
defined chunked order for CSR products); CSR products additionally within 1e-5 relative of float64 scipy."""
import math

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp
import torch

import oracle
from flyhip import hip, synth
from flyhip.brain import Brain, LIFParams
from flyhip.connectome import Connectome
from flyhip.fly import FlyBrain
from helpers import fly_setup, grouped_retina, oracle_fly, random_csr

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV)


def N(t):
    return t.detach().cpu().numpy()


def assert_bits(a, b, what=""):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if a.dtype.kind == "f":
        ok = (a == b) | (np.isnan(a) & np.isnan(b))
        if not ok.all():
            i = np.argwhere(~ok)[0]
            raise AssertionError(f"{what}: {(~ok).sum()} of {a.size} differ, first at {tuple(i)}: {a[tuple(i)]!r} vs {b[tuple(i)]!r}")
    else:
        assert np.array_equal(a, b), what


class FakeBrain:
    """The attributes lif_update reads, over arbitrary (B, N) state."""
    def __init__(self, st, flags, P, B, N):
        self.v, self.g, self.drive, self.adapt, self.refrac = (T(st[k]) for k in ("v", "g", "drive", "adapt", "refrac"))
        self.active, self.poisson_p, self.res, self.std_u_vec = (T(st[k]) for k in ("active", "poisson_p", "res", "std_u"))
        self.spikes, self.rate, self.spike_counts = T(st["spikes"]), T(st["rate"]), T(st["spike_counts"])
        self.spike_buf = T(st["out"][None])
        self.buf_pos = 0
        self._hip_P = T(P)
        self._poisson_on, self._std_on = bool(flags & 1), bool(flags & 2)
        self.record_activity = bool(flags & 4)
        self.cuda_compact = bool(flags & 32)
        self.p = LIFParams(adapt_jump=float(P[6]) if flags & 8 else 0.0)
        self.B, self.n = B, N


def random_lif_state(rng, B, N):
    st = {
        "v": rng.uniform(-60, -40, (B, N)), "g": rng.normal(0, 5, (B, N)), "drive": rng.normal(0, 3, (B, N)),
        "adapt": rng.uniform(0, 2, (B, N)), "refrac": rng.choice([0.0, 0.0, 0.1, 1.3, 2.2], (B, N)),
        "active": (rng.random(N) > 0.2), "poisson_p": np.where(rng.random((B, N)) < 0.3, rng.uniform(0, 0.5, (B, N)), 0),
        "res": rng.uniform(0.2, 1, (B, N)), "std_u": rng.uniform(0, 0.5, N), "spikes": np.zeros((B, N)),
        "out": np.zeros((B, N)), "rate": rng.uniform(0, 50, (B, N)), "spike_counts": rng.integers(0, 5, (B, N)),
    }
    st = {k: np.ascontiguousarray(v, np.float32) for k, v in st.items()}
    # frozen cells exactly at rest exercise the compact skip
    rest = rng.random((B, N)) < 0.2
    for k, val in (("v", -52.0), ("g", 0), ("drive", 0), ("adapt", 0), ("refrac", 0), ("rate", 0), ("res", 1)):
        st[k][rest] = val
    st["v"][0, :3] = -45.0                     # exactly at threshold
    return st


@pytest.mark.parametrize("B,flags", [(1, 0), (3, 1 | 4), (2, 1 | 2 | 4 | 8), (3, 1 | 32), (2, 1 | 2 | 8 | 32)])
def test_lif_update_bit_exact(B, flags):
    rng = np.random.default_rng(7 + flags)
    N = 259
    P = np.array([-52.0, -52.0, -45.0, math.exp(-0.1 / 20), 0.1, 2.2, 1.37, math.exp(-0.1 / 200),
                  math.exp(-0.1 / 300), math.exp(-0.1 / 100), (1 - math.exp(-0.1 / 100)) * 1000 / 0.1], np.float32)
    st = random_lif_state(rng, B, N)
    rnd = rng.random((B, N)).astype(np.float32)
    fb = FakeBrain(st, flags, P, B, N)
    for it in range(5):
        hip.lif_update(fb, T(rnd))
        oracle.lif(st, flags, P, rnd)
        torch.cuda.synchronize()
        for k, t in (("v", fb.v), ("g", fb.g), ("adapt", fb.adapt), ("refrac", fb.refrac), ("res", fb.res),
                     ("spikes", fb.spikes), ("rate", fb.rate), ("spike_counts", fb.spike_counts)):
            assert_bits(N_(t), st[k], f"{k} iter {it}")
        assert_bits(N_(fb.spike_buf[0]), st["out"], "out")
        rnd = rng.random((B, N)).astype(np.float32)


def N_(t):
    return t.detach().cpu().numpy()


def test_lif_phase_coefficients_bit_exact():
    rng = np.random.default_rng(3)
    B, N = 2, 300
    P = np.array([-52.0, -52.0, -45.0, 0.99, 0.1, 2.2, 1.0, 0.995, 0.997, 0.999, 10.0], np.float32)
    phase = np.stack([rng.uniform(0.9, 1, N), rng.choice([0.1, 0.2], N), rng.uniform(0.9, 1, N), rng.uniform(0.9, 1, N),
                      rng.uniform(0.99, 1, N), rng.uniform(5, 10, N), rng.choice([0.0, 1.0], N), rng.choice([1.0, 2.0], N)]
                     ).astype(np.float32)
    st = random_lif_state(rng, B, N)
    rnd = rng.random((B, N)).astype(np.float32)
    flags = 1 | 2 | 4 | 8 | 16
    fb = FakeBrain(st, flags, P, B, N)
    hip.lif_update(fb, T(rnd), T(phase))
    oracle.lif(st, flags, P, rnd, phase)
    for k, t in (("v", fb.v), ("refrac", fb.refrac), ("spikes", fb.spikes), ("rate", fb.rate), ("res", fb.res)):
        assert_bits(N_(t), st[k], k)


def test_uniform_matches_oracle():
    out = torch.empty(3, 1000, device=DEV)
    for step in (0, 1, 12345, 2**32 - 1):
        hip.uniform(out, 0xDEADBEEF12345678, step, stream_tag=1)
        assert_bits(N_(out), oracle.uniform(3, 1000, 0xDEADBEEF12345678, step, 1), f"step {step}")
    base = torch.tensor([77], dtype=torch.int32, device=DEV)
    hip.uniform(out, 5, 3, stream_tag=0, step_base=base)
    assert_bits(N_(out), oracle.uniform(3, 1000, 5, 80, 0), "step_base")


def _csc(W):
    Wc = sp.csc_matrix(W)
    Wc.sum_duplicates()
    Wc.sort_indices()
    return Wc


@pytest.mark.parametrize("B,density,activity", [(1, 0.05, 0.05), (3, 0.02, 0.3), (2, 0.2, 1.0), (1, 0.01, 0.0)])
def test_pull_and_events_bit_exact(B, density, activity):
    rng = np.random.default_rng(11)
    n = 1500
    W = random_csr(rng, n, n, density, long_rows=(3, 700), long_len=900)
    W.data *= 7.0
    x = (rng.random((B, n)) < activity).astype(np.float32)
    x[:, ::7] *= rng.uniform(0.2, 1.0, (B, len(range(0, n, 7)))).astype(np.float32)   # STD-style fractional x
    g0 = rng.normal(0, 4, (B, n)).astype(np.float32)
    a_s = math.exp(-0.1 / 5)
    Wc = _csc(W)
    ref = g0.copy()
    oracle.scale(ref, a_s)
    oracle.events(Wc.indptr, Wc.indices, Wc.data, x, ref)
    g = T(g0)
    hip.pull(T(W.indptr.astype(np.int32)), T(W.indices.astype(np.int32)), T(W.data), T(x), g, a_s)
    assert_bits(N_(g), ref, "pull")
    # the CSC-compatible scatter (no decay), full and compact pre lists
    for compact in (False, True):
        g = T(g0)
        hip.scale_(g, a_s)
        pre_ids = T(np.flatnonzero(np.diff(Wc.indptr)).astype(np.int32)) if compact else None
        hip.event_scatter(T(Wc.indptr.astype(np.int32)), T(Wc.indices.astype(np.int32)), T(Wc.data), T(x), g, pre_ids)
        assert_bits(N_(g), ref, f"events compact={compact}")


@pytest.mark.parametrize("activity", [0.02, 0.6])
def test_events_over_several_tiles_and_rounds(activity):
    """launch_events on a graph of several post tiles (10,000 cells, tiles of 4,096 posts, the last one partial), at
    an activity that needs one round of transmitting cells and at one that needs two (more than 4,096 transmitting
    cells), with fractional and negative x, posts of high in-degree (whose contributions in a round span several
    placement parts), the full and the compact walk: bit-exact against the oracle's ascending-presynaptic order."""
    rng = np.random.default_rng(17)
    n, B = 10000, 2
    W = random_csr(rng, n, n, 0.002, long_rows=(5, 4096, 9999), long_len=6000)
    x = (rng.random((B, n)) < activity).astype(np.float32)
    x[:, ::3] *= rng.uniform(-1.0, 1.0, (B, len(range(0, n, 3)))).astype(np.float32)
    g0 = rng.normal(0, 4, (B, n)).astype(np.float32)
    Wc = _csc(W)
    ref = g0.copy()
    oracle.events(Wc.indptr, Wc.indices, Wc.data, x, ref)
    for compact in (False, True):
        g = T(g0)
        pre_ids = T(np.flatnonzero(np.diff(Wc.indptr)).astype(np.int32)) if compact else None
        hip.event_scatter(T(Wc.indptr.astype(np.int32)), T(Wc.indices.astype(np.int32)), T(Wc.data), T(x), g, pre_ids)
        assert_bits(N_(g), ref, f"events compact={compact}")


def test_pull_half_weights():
    rng = np.random.default_rng(2)
    n, B = 600, 2
    W = random_csr(rng, n, n, 0.05)
    wh = W.data.astype(np.float16)
    x = (rng.random((B, n)) < 0.2).astype(np.float32)
    g0 = rng.normal(0, 1, (B, n)).astype(np.float32)
    ref = g0.copy()
    oracle.scale(ref, 0.9)
    Wc = _csc(sp.csr_matrix((wh.astype(np.float32), W.indices, W.indptr), shape=W.shape))
    oracle.events(Wc.indptr, Wc.indices, Wc.data, x, ref)
    g = T(g0)
    hip.pull(T(W.indptr.astype(np.int32)), T(W.indices.astype(np.int32)), T(wh), T(x), g, 0.9)
    assert_bits(N_(g), ref, "pull half")


@pytest.mark.parametrize("dtype", ["float32", "float16"])
def test_brain_synaptic_input_accumulates_in_float32(dtype):
    """Brain._add_synaptic_input (flyverse/brain.py:668-692): g += W x with no decay.  Half weights beyond the half
    range (6 x 20000 = 120000 > 65504) still sum exactly in float32, as flyverse/tests/test_control.py:200-207 checks;
    on a random graph the result matches the oracle's ascending-presynaptic fold bit for bit."""
    n = 6
    c = Connectome(pd.DataFrame({"bodyId": np.arange(n), "type": [f"T{i}" for i in range(n)]}),
                   sp.csr_matrix(np.ones((n, n), np.float32)))
    b = Brain(c, LIFParams(weight_dtype=dtype, input_norm_alpha=0, conn_cap=0), device=DEV)
    b.set_weights(sp.csr_matrix(np.full((n, n), 20000, dtype=np.float32)))
    b._add_synaptic_input(torch.ones_like(b.g))
    assert b.g.dtype == torch.float32
    assert torch.equal(b.g, torch.full_like(b.g, 120000.0))
    rng = np.random.default_rng(8)
    n, B = 500, 3
    W = random_csr(rng, n, n, 0.05)
    if dtype == "float16":
        W.data = W.data.astype(np.float16).astype(np.float32)
    c = Connectome(pd.DataFrame({"bodyId": np.arange(n), "type": ["T"] * n}), sp.csr_matrix((n, n), dtype=np.float32))
    b = Brain(c, LIFParams(weight_dtype=dtype), device=DEV, batch=B)
    b.set_weights(W)
    g0 = rng.normal(0, 1, (B, n)).astype(np.float32)
    x = (rng.random((B, n)) < 0.3).astype(np.float32)
    b.g.copy_(T(g0))
    b._add_synaptic_input(T(x))
    ref = g0.copy()
    oracle.events(_csc(W).indptr, _csc(W).indices, _csc(W).data, x, ref)
    assert_bits(N_(b.g), ref, f"synaptic input {dtype}")
    with pytest.raises(ValueError):
        b._add_synaptic_input(T(x[:, :-1]))


@pytest.mark.parametrize("B,kernel,range_cols", [(1, "grouped", 0), (3, "grouped", 0), (8, "grouped", 0),
                                                  (1, "chunked", 0), (3, "chunked", 0), (1, "lds", 0),
                                                  (3, "lds", 1024), (2, "lds", 1500), (1, "panel", 0),
                                                  (3, "panel", 512), (2, "panel", 1500), (3, "panel", 999),
                                                  (2, "panel", 333), (1, "panel", 64)])
def test_csr_products_bit_exact_and_close_to_float64(B, kernel, range_cols):
    rng = np.random.default_rng(5 + B)
    rows, cols = 2000, 1700
    M = random_csr(rng, rows, cols, 0.03, long_rows=(0, 999, 1998), long_len=1650)
    # empty rows in the middle and at the end, a block of rows of exactly 16 and 17 entries
    dense = sp.csr_matrix(np.where(np.arange(cols)[None, :] < np.array([16, 17, 15, 1])[:, None],
                                   rng.standard_normal((4, cols)), 0).astype(np.float32))
    M = sp.vstack([M[:700], sp.csr_matrix((2, cols), dtype=np.float32), dense, M[700:],
                   sp.csr_matrix((3, cols), dtype=np.float32)]).tocsr()
    M.sort_indices()
    rows = M.shape[0]
    x = rng.standard_normal((B, cols)).astype(np.float32)
    x[:, ::3] = 0
    ref = oracle.spmv(M.indptr, M.indices, M.data, x)
    csr = hip.CSR(M, DEV, kernel=kernel, range_cols=range_cols)
    assert csr.kernel == kernel
    y = csr.matvec(T(x))
    assert_bits(N_(y), ref, kernel)
    acc = T(rng.standard_normal((B, rows)).astype(np.float32))
    want = (N_(acc) + (np.float32(0.5) * ref).astype(np.float32)).astype(np.float32)
    assert_bits(N_(csr.matvec(T(x), out=acc, scale=0.5, accumulate=True)), want, "scale + accumulate")
    assert_bits(N_(csr.matvec_reference_shape(T(x))), ref, "launch_csr")
    exact = (M.astype(np.float64) @ x.astype(np.float64).T).T
    np.testing.assert_allclose(N_(y), exact, rtol=1e-5, atol=1e-5 * np.abs(exact).max())
    y2 = csr.matvec(T(x), scale=3.0)
    assert_bits(N_(y2), (np.float32(3.0) * ref).astype(np.float32), "scaled")
    # float16 weights: x rounded to half first (the reference warp kernel's convention)
    h = hip.CSR(M, DEV, torch.float16)
    wh = M.data.astype(np.float16).astype(np.float32)
    assert_bits(N_(h.matvec(T(x))), oracle.spmv(M.indptr, M.indices, wh, x, half=True), "half")


def test_copy_probe_copies_and_refuses_misaligned_buffers():
    """The bandwidth probe bench.py times (launch_copy_stream): an exact copy of 4 * n4 floats at any grid, the
    floats past them untouched, and an argument error (nothing launched) for a buffer that is not 16-byte aligned."""
    lib = hip.lib(DEV)
    stream = torch.cuda.current_stream().cuda_stream
    x = torch.randn(4 * 1000 + 8, device=DEV)
    for grid in (1, 7, 256):
        y = torch.full_like(x, -1.0)
        assert lib.launch_copy_stream(x.data_ptr(), y.data_ptr(), 1000, grid, stream) == 0
        torch.cuda.synchronize()
        assert torch.equal(y[:4000], x[:4000]) and bool((y[4000:] == -1.0).all())
    y = torch.zeros_like(x)
    assert lib.launch_copy_stream(x[1:].data_ptr(), y.data_ptr(), 10, 4, stream) < 0
    assert lib.launch_copy_stream(x.data_ptr(), y.data_ptr(), 10, 0, stream) < 0
    assert lib.launch_copy_stream(x.data_ptr(), y.data_ptr(), 0, 4, stream) == 0
    torch.cuda.synchronize()
    assert not bool(y.any())


@pytest.mark.parametrize("B,half", [(1, False), (3, False), (2, True)])
def test_reference_csr_product_over_row_windows(B, half):
    """launch_csr (the neural.cu:105 contract) on rows of every length class of its windows of four rows: empty,
    shorter than a chunk, one chunk exactly, just over, the 256-entry bound of a shared chunk stream and just over it
    (the window's rows then run one at a time), in every position of a window, and a last partial window; bit-exact
    against the oracle's chunked order (x rounded to half first for half weights)."""
    rng = np.random.default_rng(31 + B)
    cols = 700
    lengths = [0, 1, 3, 16, 17, 64, 65, 255, 256, 257, 300, 2, 15, 0, 33, 48, 49, 256, 1, 5, 9]
    lengths = lengths + lengths[5:] + lengths[:4]                  # 41 rows: the last window holds one
    rows = []
    for L in lengths:
        c = np.sort(rng.choice(cols, L, replace=False))
        rows.append(sp.csr_matrix((rng.standard_normal(L).astype(np.float32), c, [0, L]), shape=(1, cols)))
    M = sp.vstack(rows).tocsr()
    if half:
        M.data = M.data.astype(np.float16).astype(np.float32)
    x = rng.standard_normal((B, cols)).astype(np.float32)
    x[:, ::5] = 0
    x[:, 1::11] = -0.0
    csr = hip.CSR(M, DEV, torch.float16 if half else torch.float32)
    ref = oracle.spmv(M.indptr, M.indices, M.data, x, half=half)
    assert_bits(N_(csr.matvec_reference_shape(T(x))), ref, "launch_csr")


def test_csr_empty_and_validation():
    csr = hip.CSR(sp.csr_matrix((3, 4), dtype=np.float32), DEV)
    x = torch.zeros(2, 4, device=DEV)
    assert torch.equal(csr.matvec(x), torch.zeros(2, 3, device=DEV))
    b = torch.zeros(4, device=DEV)
    with pytest.raises(ValueError):
        hip.optic_dr(x.double(), b, x)
    with pytest.raises(ValueError):
        hip.optic_dr(x[:, ::2], b[:2], x[:, ::2])
    with pytest.raises(ValueError):
        hip.optic_dr(x, b[:2], x)
    v = torch.empty(2, 0, device=DEV)
    hip.optic_dr(v, b[:0], v)


def test_optic_kernels_bit_exact_on_a_side_stream():
    from types import SimpleNamespace
    from flyhip.optic import OpticParams
    stream = torch.cuda.Stream()
    rng = np.random.default_rng(6)
    with torch.cuda.stream(stream):
        for B in (1, 3):
            n = 259
            v = (rng.random((B, n)) - 0.5).astype(np.float32)
            adapt = rng.random((B, n)).astype(np.float32)
            a = rng.random(n).astype(np.float32)
            b = rng.random(n).astype(np.float32)
            o = SimpleNamespace(v=T(v), adapt=T(adapt), B=B, n_rate=n, p=OpticParams(), _a=T(a), b_vec=T(b),
                                _a_ad=0.997, _cuda_dr=torch.empty(B, n, device=DEV))
            hip.optic_dr(o.v, o.b_vec, o._cuda_dr)
            dr = oracle.optic_dr(v, b)
            for _ in range(10):
                y, pr, sk = (rng.random((B, n)).astype(np.float32) for _ in range(3))
                hip.optic_update(o, T(y), T(pr), T(sk))
                oracle.optic_update(v, adapt, dr, y, pr, sk, a, b, 1.0, 1.0, 0.997)
            stream.synchronize()
            assert_bits(N_(o.v), v, "v")
            assert_bits(N_(o.adapt), adapt, "adapt")
            assert_bits(N_(o._cuda_dr), dr, "dr")


@pytest.mark.parametrize("grouped", [False, True])
def test_photoreceptor_feedback_drive_bit_exact(grouped):
    """The photoreceptor stage against the oracle: one photoreceptor per column, or several per (column, family)
    cell with empty cells and columns (the averaging matrix of flyverse/optic.py:320-325)."""
    from flyhip.connectome import Retina
    rng = np.random.default_rng(8)
    B, n_pr = 3, 777
    sens = rng.random((n_pr, 4)).astype(np.float32)
    if grouped:
        retina = Retina(np.arange(n_pr), sens, rng.integers(0, 200, n_pr), rng.choice(5, n_pr, p=[.6, .1, .1, .1, .1]),
                        205)
    else:
        retina = Retina(np.arange(n_pr), sens)
    cells = retina.cells()
    tcells = {k: T(v) for k, v in cells.items()}
    n_col, n_cells = retina.n_col, retina.n_cells
    coefs = np.array([math.exp(-1), 1 - math.exp(-1), math.exp(-10 / 300), 1 - math.exp(-10 / 300), 0.02, 2.0], np.float32)
    I_lp, I_mean = np.zeros((B, n_cells), np.float32), np.zeros((B, n_cells), np.float32)
    fresh = np.ones(B, np.uint8)
    tI_lp, tI_mean, tfresh = T(I_lp), T(I_mean), T(fresh)
    a_pr = torch.empty(B, n_pr, device=DEV)
    contrast = torch.empty(B, n_cells, device=DEV)
    for _ in range(4):
        rad = rng.random((B, n_col, 4)).astype(np.float32)
        hip.photoreceptor(T(rad), T(sens), tcells, tI_lp, tI_mean, tfresh, T(coefs), contrast, a_pr)
        con = np.empty((B, n_cells), np.float32)
        ref = oracle.photoreceptor(rad, sens, I_lp, I_mean, fresh, coefs, cells, contrast=con)
        assert_bits(N_(a_pr), ref, "a_pr")
        assert_bits(N_(contrast), con, "contrast")
        assert_bits(N_(tI_mean), I_mean, "I_mean")
        assert_bits(N_(tI_lp), I_lp, "I_lp")
    assert int(tfresh.sum()) == 0
    with pytest.raises(ValueError):
        hip.photoreceptor(T(rad[:, :-1]), T(sens), tcells, tI_lp, tI_mean, tfresh, T(coefs), contrast, a_pr)
    n = 900
    rate = rng.uniform(0, 400, (B, n)).astype(np.float32)
    spk = np.sort(rng.choice(n, 300, replace=False)).astype(np.int32)
    s = torch.empty(B, 300, device=DEV)
    hip.feedback(T(rate), T(spk), s)
    assert_bits(N_(s), oracle.feedback(rate, spk), "feedback")
    y = rng.normal(0, 0.5, (B, 300)).astype(np.float32)
    drive = torch.zeros(B, n, device=DEV)
    hip.drive_out(T(y), T(spk), drive, 100.0, 35.0)
    assert_bits(N_(drive), oracle.drive_out(y, spk, n, 100.0, 35.0), "drive")


def test_group_means():
    rng = np.random.default_rng(9)
    rates = rng.uniform(0, 100, (3, 50)).astype(np.float32)
    ptr = np.array([0, 3, 3, 10], np.int32)
    idx = rng.integers(0, 50, 10).astype(np.int32)
    out = hip.group_means(T(ptr), T(idx), T(rates), T(np.ones((3, 50), np.float32)))
    ref = oracle.group_means(ptr, idx, rates)
    assert np.array_equal(N_(out)[:, :3], ref)
    assert np.array_equal(N_(out)[:, 3], np.full(3, 50.0))


def test_brain_group_rates_and_spike_total_snapshot():
    """Brain.group_rates (flyverse/motor.py:232-269): the fixed-order means equal the oracle's, an empty group reads
    0, and total_spikes reuses the copied spike total until a step or a reset changes the state."""
    c, inputs, hz = synth.plain_lif_graph(5, n=2000, n_inputs=200)
    brain = Brain(c, LIFParams.shiu(dt=0.1), device=DEV, batch=3, seed=7)
    brain.set_poisson(inputs, hz)
    brain.step(60)
    rng = np.random.default_rng(4)
    groups = [rng.integers(0, c.n, 17), np.zeros(0, np.int64), inputs, np.arange(c.n)]
    with pytest.raises(ValueError):
        brain.group_rates([np.array([c.n])])
    out = brain.group_rates(groups)
    ptr = np.r_[0, np.cumsum([len(g) for g in groups])].astype(np.int32)
    ref = oracle.group_means(ptr, np.concatenate(groups).astype(np.int32), N_(brain.rate))
    assert out.shape == (3, 4) and np.array_equal(out, ref) and np.all(out[:, 1] == 0.0)
    assert out[:, 2].min() > 0
    spikes = N_(brain.spikes).sum(1)
    assert np.array_equal(brain.total_spikes(), spikes)
    brain._spike_total[1][:] = -1.0                       # the snapshot, not a fresh sum, answers until the state moves
    assert np.all(brain.total_spikes() == -1.0)
    brain.step()
    assert np.array_equal(brain.total_spikes(), N_(brain.spikes).sum(1))
    assert np.array_equal(brain.group_rates(groups[:1]), oracle.group_means(ptr[:2], groups[0].astype(np.int32),
                                                                            N_(brain.rate)))
    brain.reset()
    assert np.all(brain.total_spikes() == 0.0)


def test_brain_raster_bit_exact_plain_lif():
    """The plain-LIF workload's statistics at 3,000 neurons: Brain on the GPU against the oracle, 400 steps."""
    c, inputs, hz = synth.plain_lif_graph(4, n=3000, n_inputs=60)
    p = LIFParams.shiu(dt=0.1)
    for B in (1, 2):
        brain = Brain(c, p, device=DEV, batch=B, seed=123)
        brain.record_activity = True
        brain.set_poisson(inputs, hz)
        from helpers import oracle_fly
        om = oracle_fly(c, None, p, None, B, 123, None, with_optic=False)
        om.st.poisson_p[:, inputs] = (hz * np.float32(p.dt / 1000.0)).astype(np.float32)
        total = 0
        for step, st in enumerate(om.lif_steps(400)):
            brain.step()
            s = N_(brain.spikes)
            assert_bits(s, st.spikes, f"spikes B={B} step {step}")
            total += s.sum()
        for k in ("v", "g", "refrac", "rate", "spike_counts"):
            assert_bits(N_(getattr(brain, k)), getattr(om.st, k), f"{k} B={B}")
        assert total > 0


def test_optic_lobe_photoreceptor_activity_and_type_summary():
    """OpticLobe.photoreceptor_activity (flyverse/optic.py:496-517) on a grouped retina against the oracle, and the
    per-type summary of the last frame's rate change (optic.py:696-703)."""
    from flyhip.optic import OpticLobe, photoreceptor_coefs
    conn, retina, layout, lp, op = fly_setup(0.02, seed=4)
    retina = grouped_retina(retina, per_column=3, seed=5)
    B = 2
    ol = OpticLobe(conn, retina, op, device=DEV, batch=B)
    assert ol.delta_rate_by_type() is None
    cells = retina.cells()
    I_lp = np.zeros((B, retina.n_cells), np.float32)
    I_mean = np.zeros_like(I_lp)
    fresh = np.ones(B, np.uint8)
    rng = np.random.default_rng(6)
    for frame in range(3):
        rad = rng.random((B, retina.n_col, 4)).astype(np.float32)
        got = ol.photoreceptor_activity(T(rad), 10.0)
        ref = oracle.photoreceptor(rad, retina.pr_sens, I_lp, I_mean, fresh, photoreceptor_coefs(op, 10.0), cells)
        assert_bits(N_(got), ref, f"activity frame {frame}")
    assert_bits(N_(ol.I_mean), I_mean, "I_mean")
    ol.step_frame(T(rad[0]), torch.zeros(B, conn.n, device=DEV), 10.0)
    summary = ol.delta_rate_by_type(top=5, row=1)
    assert list(summary.columns) == ["mean_dr", "mean_abs", "n"] and 0 < len(summary) <= 5
    assert summary["mean_abs"].is_monotonic_decreasing and summary["n"].sum() <= ol.n_rate


@pytest.mark.parametrize("graphs", [False, True])
def test_fly_frames_bit_exact_with_optic(graphs):
    """A 2 % scale fly (LIF + optic lobe + photoreceptor movie): 5 frames of 10 ms, rasters identical step for step.
    The graph-captured run places its photoreceptors several to a (column, family) cell."""
    conn, retina, layout, lp, op = fly_setup(0.02, seed=1)
    if graphs:
        retina = grouped_retina(retina, per_column=4, seed=3)
    B = 2
    fb = FlyBrain(conn, retina, batch=B, device=DEV, seed=9, lif_params=lp, optic_params=op, cuda_graphs=graphs)
    idx, hz = synth.fly_inputs(5, layout, 40)
    fb.set_poisson(idx, hz)
    fb.random_movie(17)
    om = oracle_fly(conn, retina, lp, op, B, 9, 17)
    om.st.poisson_p[:, idx] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    for frame in range(5):
        raster = []
        om.frame(100, raster)
        fb.step(10.0)
        assert_bits(N_(fb.brain.spikes), raster[-1], f"spikes frame {frame}")
        assert_bits(N_(fb.brain.drive), om.st.drive, f"drive frame {frame}")
        assert_bits(N_(fb.optic.v), om.os.v, f"optic v frame {frame}")
        assert_bits(N_(fb.optic.I_lp), om.os.I_lp, f"photoreceptor low-pass frame {frame}")
        assert_bits(N_(fb.brain.spike_counts), om.st.spike_counts, f"counts frame {frame}")
    for k in ("v", "g", "rate", "refrac"):
        assert_bits(N_(getattr(fb.brain, k)), getattr(om.st, k), k)
    assert om.st.spike_counts.sum() > 0


@pytest.mark.parametrize("graphs", [False, True])
def test_fly_state_dict_round_trip(graphs):
    """state_dict (the reference's tensor names, flyverse/fly.py:721-792) mid-run, then load_state_dict into a fresh
    FlyBrain: both continue with the same bits as the uninterrupted run."""
    conn, retina, layout, lp, op = fly_setup(0.02, seed=2)
    idx, hz = synth.fly_inputs(6, layout, 40)

    def make():
        fb = FlyBrain(conn, retina, batch=2, device=DEV, seed=4, lif_params=lp, optic_params=op, cuda_graphs=graphs)
        fb.set_poisson(idx, hz)
        fb.random_movie(23)
        return fb

    a = make()
    for _ in range(3):
        a.step(10.0)
    state = a.state_dict()
    assert set(state["brain"]) == set(FlyBrain.BRAIN_TENSORS)
    for _ in range(2):
        a.step(10.0)
    b = make()
    b.step(10.0)                         # some other history first, then the checkpoint replaces it
    b.load_state_dict(state)
    for _ in range(2):
        b.step(10.0)
    for k in ("v", "g", "rate", "refrac", "spikes", "spike_counts", "drive"):
        assert torch.equal(getattr(a.brain, k), getattr(b.brain, k)), k
    for k in ("v", "adapt", "delta_rate"):
        assert torch.equal(getattr(a.optic, k), getattr(b.optic, k)), k
    assert a.brain.spike_counts.sum().item() > 0
    assert a.t == b.t == a.brain.t
    other = dict(state, body_ids=state["body_ids"] + 1)
    with pytest.raises(ValueError, match="connectome does not match"):
        b.load_state_dict(other)
    # another seed: the checkpoint's Poisson key comes with the state, so the run continues as the original
    c = FlyBrain(conn, retina, batch=2, device=DEV, seed=99, lif_params=lp, optic_params=op, cuda_graphs=graphs)
    c.set_poisson(idx, hz)
    c.random_movie(23)
    c.load_state_dict(state)
    assert c.brain.seed == 4
    for _ in range(2):
        c.step(10.0)
    for k in ("v", "g", "spikes", "spike_counts"):
        assert torch.equal(getattr(a.brain, k), getattr(c.brain, k)), k
    # other dynamics: refused
    from dataclasses import replace
    d = FlyBrain(conn, retina, batch=2, device=DEV, seed=4, lif_params=replace(lp, tau_m=lp.tau_m * 2),
                 optic_params=op, cuda_graphs=graphs)
    with pytest.raises(ValueError, match="simulation parameters do not match"):
        d.load_state_dict(state)
    e = FlyBrain(conn, retina, batch=2, device=DEV, seed=4, lif_params=lp, optic_params=replace(op, gain_rr=0.5),
                 cuda_graphs=graphs)
    with pytest.raises(ValueError, match="simulation parameters do not match"):
        e.load_state_dict(state)


@pytest.mark.parametrize("graphs", [False, True])
def test_fly_reset_starts_over(graphs):
    """reset() (flyverse/fly.py:682-706) returns a fly to its built state: clocks at zero, no held radiance, no
    captured graphs, no Poisson drive; given the same inputs again it gives the bits of a fresh FlyBrain.  reset(rows)
    rests those flies only."""
    conn, retina, layout, lp, op = fly_setup(0.02, seed=2)
    idx, hz = synth.fly_inputs(6, layout, 40)

    def make():
        return FlyBrain(conn, retina, batch=2, device=DEV, seed=4, lif_params=lp, optic_params=op, cuda_graphs=graphs)

    def run(fb):
        fb.set_poisson(idx, hz)
        fb.random_movie(23)
        for _ in range(3):
            fb.step(10.0)

    a = make()
    run(a)
    a.step(2.5)                                      # leaves a fractional carry behind
    a.reset()
    assert a.t == 0.0 and a.brain.step_count == 0 and a.brain.buf_pos == 0 and a._pending_ms == 0.0
    assert a._radiance is None and not a._graphs and not a.brain._poisson_on
    assert not bool(a.brain.poisson_p.any())
    run(a)
    b = make()
    run(b)
    for k in ("v", "g", "rate", "refrac", "spikes", "spike_counts", "drive", "spike_buf"):
        assert torch.equal(getattr(a.brain, k), getattr(b.brain, k)), k
    for k in ("v", "adapt", "I_lp", "delta_rate"):
        assert torch.equal(getattr(a.optic, k), getattr(b.optic, k)), k
    assert a.t == b.t and a.brain.spike_counts.sum().item() > 0
    a.reset([1])
    assert torch.equal(a.brain.v[0], b.brain.v[0]) and bool((a.brain.v[1] == lp.v_rest).all())
    assert not bool(a.brain.spike_buf[:, 1].any()) and not bool(a.brain.poisson_p[1].any())
    assert a.t == b.t


def test_fly_graphs_follow_weight_replacement():
    """set_weights invalidates the captured frames (flyverse/fly.py:513-515, test_cuda.py:163-167): after new
    weights the graph path captures afresh and keeps the bits of the eager path."""
    conn, retina, layout, lp, op = fly_setup(0.02, seed=2)
    idx, hz = synth.fly_inputs(6, layout, 40)
    flies = []
    for graphs in (False, True):
        fb = FlyBrain(conn, retina, batch=2, device=DEV, seed=4, lif_params=lp, optic_params=op, cuda_graphs=graphs)
        fb.set_poisson(idx, hz)
        fb.random_movie(23)
        fb.step(10.0)
        fb.step(10.0)
        flies.append(fb)
    eager, graphed = flies
    old = list(graphed._graphs.values())
    assert old
    for fb in flies:
        fb.brain.set_weights(fb.brain._W_cpu * np.float32(1.5))
        fb.step(10.0)
        fb.step(10.0)
    assert all(entry not in old for entry in graphed._graphs.values())
    for k in ("v", "g", "rate", "spikes", "spike_counts", "drive"):
        assert torch.equal(getattr(eager.brain, k), getattr(graphed.brain, k)), k
    assert eager.brain.spike_counts.sum().item() > 0


@pytest.mark.parametrize("graphs", [False, True])
def test_fly_step_budget(graphs):
    """step_budget (flyverse/fly.py:618-660): argument checks, the StepResult accounting, and the chunks it took
    replayed as plain step() calls on a twin brain give the same bits."""
    conn, retina, layout, lp, op = fly_setup(0.02, seed=3)
    idx, hz = synth.fly_inputs(6, layout, 40)

    def make():
        fb = FlyBrain(conn, retina, batch=2, device=DEV, seed=5, lif_params=lp, optic_params=op, cuda_graphs=graphs)
        fb.set_poisson(idx, hz)
        fb.random_movie(29)
        return fb

    a = make()
    for bad in (dict(wall_ms=-1.0), dict(wall_ms=float("nan")), dict(wall_ms=5.0, frame_ms=0.05)):
        with pytest.raises(ValueError):
            a.step_budget(**bad)
    chunks = []
    total = 0
    for wall in (40.0, 0.0, 60.0):
        before = a.brain.step_count
        r = a.step_budget(wall, frame_ms=2.0)
        chunks += a._budget_chunks
        assert r.steps == a.brain.step_count - before == sum(a._budget_chunks)
        assert r.simulated_ms == pytest.approx(r.steps * lp.dt, abs=1e-9)
        assert r.wall_ms >= 0 and (r.steps > 0) == (wall > 0)
        assert all(1 <= n <= 20 for n in a._budget_chunks)
        total += r.steps
    b = make()
    for n in chunks:
        b.step(n * lp.dt)
    assert b.brain.step_count == a.brain.step_count == total
    for k in ("v", "g", "rate", "refrac", "spikes", "spike_counts", "drive"):
        assert torch.equal(getattr(a.brain, k), getattr(b.brain, k)), k
    for k in ("v", "adapt", "delta_rate"):
        assert torch.equal(getattr(a.optic, k), getattr(b.optic, k)), k


def _pack_bits(x):
    B, n = x.shape
    nw = (n + 63) // 64
    on = np.zeros((B, nw * 64), np.uint64)
    on[:, :n] = (x != 0)
    w = (on.reshape(B, nw, 64) << np.arange(64, dtype=np.uint64)).sum(-1, dtype=np.uint64)
    return w.view(np.int64)


@pytest.mark.parametrize("B,activity", [(1, 0.02), (3, 0.3)])
def test_pull_bits_bit_exact(B, activity):
    rng = np.random.default_rng(21)
    n = 2100
    W = random_csr(rng, n, n, 0.03, long_rows=(5,), long_len=1500)
    W = sp.vstack([W[:1900], sp.csr_matrix((200, n), dtype=np.float32)]).tocsr()   # 200 rows without in-edges
    W.data *= 5.0
    x = (rng.random((B, n)) < activity).astype(np.float32)
    g0 = rng.normal(0, 3, (B, n)).astype(np.float32)
    Wc = _csc(W)
    ref = g0.copy()
    oracle.scale(ref, 0.98)
    oracle.events(Wc.indptr, Wc.indices, Wc.data, x, ref)
    deg = np.diff(W.indptr)
    sched = np.argsort(-deg, kind="stable").astype(np.int32)
    g = T(g0)
    hip.pull_bits(T(W.indptr.astype(np.int32)), T(W.indices.astype(np.int32)), T(W.data), T(_pack_bits(x)), g,
                  T(sched), int((deg > 0).sum()), 0.98, grid=64)
    assert_bits(N_(g), ref, "pull_bits")


def test_brain_with_depression_and_adaptation_bit_exact():
    """STD and adaptation on (the float spike_buf pull path) against the oracle's brain_step, 300 steps."""
    c, inputs, hz = synth.plain_lif_graph(6, n=2000, n_inputs=80)
    p = LIFParams.shiu(dt=0.1, std_u=0.3, adapt_jump=1.2)
    B = 2
    brain = Brain(c, p, device=DEV, batch=B, seed=77)
    brain.record_activity = True
    brain.set_poisson(inputs, hz)
    from helpers import oracle_fly
    om = oracle_fly(c, None, p, None, B, 77, None, with_optic=False)
    om.flags = 32 | 2 | 8
    om.P = oracle.lif_params_vector(p)
    om.st.std_u[:] = np.float32(p.std_u)
    om.st.poisson_p[:, inputs] = (hz * np.float32(p.dt / 1000.0)).astype(np.float32)
    for step, st in enumerate(om.lif_steps(300)):
        brain.step()
        assert_bits(N_(brain.spikes), st.spikes, f"spikes step {step}")
    for k in ("v", "g", "refrac", "rate", "adapt", "res"):
        assert_bits(N_(getattr(brain, k)), getattr(om.st, k), k)
    assert_bits(N_(brain.spike_buf), om.st.spike_buf, "spike_buf")
    assert om.st.spike_counts.sum() > 0


@pytest.mark.parametrize("B,adapt", [(1, 0.0), (2, 1.5)])
def test_brain_clocks_bit_exact(B, adapt):
    """Brain.set_clocks (flyverse/brain.py:618-646) with clocks of 1, 2, 3 and 6 steps: the clocked step on the HIP
    kernels (launch_clock_acc, launch_pull_clocked, launch_lif_rng with the phase rows, flag 16) against the oracle's
    clocked composite, every step: spikes, state and the clock accumulators."""
    c, inputs, hz = synth.plain_lif_graph(7, n=3000, n_inputs=150, hz=400.0)
    lp = LIFParams.shiu(dt=0.1, adapt_jump=adapt)
    k = np.random.default_rng(3).choice([1, 2, 3, 6], c.n)
    brain = Brain(c, lp, device=DEV, batch=B, seed=77)
    brain.record_activity = True
    brain.set_poisson(inputs, hz)
    brain.set_clocks(k)
    assert brain.K == 6 and sorted(brain._acc) == [2, 3, 6]
    om = oracle_fly(c, None, lp, None, B, 77, None, with_optic=False)
    om.flags = 32 | (8 if adapt > 0 else 0)
    om.st.poisson_p[:, inputs] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    om.set_clocks(k, om.W, lp)
    total = 0
    for step, st in enumerate(om.lif_steps(150)):
        brain.step()
        assert_bits(N(brain.spikes), st.spikes, f"spikes step {step}")
        total += int(st.spikes.sum())
        if step % 10 == 9:
            for name in ("v", "g", "refrac", "rate", "spike_counts", "adapt"):
                assert_bits(N(getattr(brain, name)), getattr(st, name), f"{name} step {step}")
            for kk, acc in om.clocks["acc"].items():
                assert_bits(N(brain._acc[kk]), acc, f"accumulator {kk} step {step}")
    assert total > 100
    with pytest.raises(ValueError, match="slowest clock"):
        brain.set_clocks(np.full(c.n, 40))


@pytest.mark.parametrize("graphs", [False, True])
def test_fly_dt_by_module_bit_exact(graphs):
    """FlyBrain with LIFParams.dt_by_module (flyverse/fly.py:98-107): the central cells on a 0.2 ms clock, the rest
    on 0.1 ms; frames with the optic lobe against the oracle's clocked composite, captured frames included."""
    from dataclasses import replace
    from flyhip.connectome import clock_multipliers
    conn, retina, layout, lp, op = fly_setup(0.02, seed=3)
    lp = replace(lp, dt_by_module={"central": 0.2})
    k = clock_multipliers(conn, lp.dt_by_module, lp.dt)
    assert set(np.unique(k)) == {1, 2}
    idx, hz = synth.fly_inputs(8, layout, 40)
    fb = FlyBrain(conn, retina, batch=2, device=DEV, seed=5, lif_params=lp, optic_params=op, cuda_graphs=graphs)
    fb.set_poisson(idx, hz)
    fb.random_movie(9)
    assert fb.brain.K == 2
    om = oracle_fly(conn, retina, lp, op, 2, 5, 9)
    om.st.poisson_p[:, idx] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    om.set_clocks(k, om.W, lp)
    for frame in range(4):
        raster = []
        om.frame(100, raster)
        fb.step(10.0)
        torch.cuda.synchronize()
        assert_bits(N(fb.brain.spikes), raster[-1], f"spikes frame {frame}")
        for name in ("v", "g", "rate", "spike_counts"):
            assert_bits(N(getattr(fb.brain, name)), getattr(om.st, name), f"{name} frame {frame}")
        assert_bits(N(fb.optic.v), om.os.v, f"optic v frame {frame}")
    assert om.st.spike_counts.sum() > 0
    # a checkpoint carries the clocks and their accumulators: a clocked twin continues with the same bits, an
    # unclocked brain refuses it (flyverse/fly.py:762-768)
    state = fb.state_dict()
    twin = FlyBrain(conn, retina, batch=2, device=DEV, seed=5, lif_params=lp, optic_params=op, cuda_graphs=graphs)
    twin.set_poisson(idx, hz)
    twin.load_state_dict(state)
    fb.step(10.0)
    twin.step(10.0)
    for name in ("v", "g", "spike_counts"):
        assert torch.equal(getattr(fb.brain, name), getattr(twin.brain, name)), name
    plain = FlyBrain(conn, retina, batch=2, device=DEV, seed=5, lif_params=replace(lp, dt_by_module=None),
                     optic_params=op)
    with pytest.raises(ValueError, match="do not match"):
        plain.load_state_dict(state)
    with pytest.raises(ValueError, match="unknown module"):
        clock_multipliers(conn, {"nowhere": 0.2}, lp.dt)
    with pytest.raises(ValueError, match="positive multiple"):
        clock_multipliers(conn, {"central": 0.15}, lp.dt)
