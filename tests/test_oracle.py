"""Pins the CPU oracle against what does not come from the oracle itself: published known-answer vectors, the
reference's own torch arithmetic (its CPU path IS torch's CPU kernels, flyverse/brain.py:838-868 and
flyverse/optic.py:505-551, so running those ops in that order is the reference computation), closed forms, brute
force in float64 and numpy's IEEE half conversion."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle
from flyhip.connectome import Retina
from helpers import random_csr


# ------------------------------------------------------------------ Philox4x32-10
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, expect):
    """Random123's published kat_vectors for philox4x32-10."""
    assert list(oracle.philox4x32_10(ctr, key)) == expect


def test_uniform_is_top_24_bits_of_the_first_word():
    seed, step = 0x0123456789ABCDEF, 77
    u = oracle.uniform(2, 50, seed, step, 3)
    for b in range(2):
        for i in (0, 7, 49):
            w = oracle.philox4x32_10([i, step, b, 3], [seed & 0xFFFFFFFF, seed >> 32])[0]
            assert u[b, i] == np.float32((int(w) >> 8) * 2.0**-24)
    big = oracle.uniform(1, 200_000, 1, 0)
    assert 0.0 <= big.min() and big.max() < 1.0
    assert abs(big.mean() - 0.5) < 0.005 and abs(big.var() - 1 / 12) < 0.002


# ------------------------------------------------------------------ LIF
def P_vec(dt=0.1, adapt=0.0, std=False):
    a_m, a_r = math.exp(-dt / 20), math.exp(-dt / 100)
    return np.array([-52, -52, -45, a_m, dt, 2.2, adapt, math.exp(-dt / 200) if adapt else 0.0,
                     math.exp(-dt / 300) if std else 1.0, a_r, (1 - a_r) * 1000 / dt], np.float32)


def torch_reference_step(s, P, rnd, flags, dt):
    """flyverse/brain.py:838-868 written with the same torch CPU ops in the same order (no phase)."""
    v, g, drive, adapt, refrac = s["v"], s["g"], s["drive"], s["adapt"], s["refrac"]
    p_rest, p_reset, p_th, t_ref, jump = float(P[0]), float(P[1]), float(P[2]), float(P[5]), float(P[6])
    a_m = math.exp(-dt / 20)
    a_r = math.exp(-dt / 100)
    target = p_rest + g + drive - adapt
    torch.add(target, (v - target) * a_m, out=v)
    in_ref = refrac > 0
    torch.where(in_ref, torch.full_like(v, p_reset), v, out=v)
    torch.clamp(refrac - dt, min=0.0, out=refrac)
    spikes = torch.mul((v >= p_th).float(), s["active"])
    if flags & 1:
        forced = (rnd < s["poisson_p"]).float()
        torch.maximum(spikes, forced, out=spikes)
    fired = spikes > 0
    torch.where(fired, torch.full_like(v, p_reset), v, out=v)
    torch.where(fired, torch.full_like(refrac, t_ref), refrac, out=refrac)
    if flags & 8:
        adapt.mul_(math.exp(-dt / 200)).add_(spikes, alpha=jump)
    if flags & 2:
        out = spikes * s["res"]
        torch.where(fired, s["res"] * (1 - s["std_u"]), s["res"], out=s["res"])
        torch.sub(1.0, (1.0 - s["res"]) * math.exp(-dt / 300), out=s["res"])
    else:
        out = spikes.clone()
    if flags & 4:
        s["spike_counts"].add_(spikes)
    s["rate"].mul_(a_r).add_(spikes * ((1 - a_r) * 1000.0 / dt))
    s["spikes"] = spikes
    s["out"] = out


@pytest.mark.parametrize("flags", [0, 1, 1 | 4, 1 | 2 | 4, 1 | 8, 1 | 2 | 4 | 8])
def test_lif_matches_reference_torch_ops(flags):
    rng = np.random.default_rng(flags)
    B, N, dt = 3, 400, 0.1
    P = P_vec(dt, adapt=1.37 if flags & 8 else 0.0, std=bool(flags & 2))
    st = {"v": rng.uniform(-60, -40, (B, N)), "g": rng.normal(0, 5, (B, N)), "drive": rng.normal(0, 3, (B, N)),
          "adapt": rng.uniform(0, 2, (B, N)), "refrac": rng.choice([0, 0, 0.1, 1.3, 2.2], (B, N)),
          "active": (rng.random(N) > 0.2), "poisson_p": np.where(rng.random((B, N)) < .3, rng.uniform(0, .5, (B, N)), 0),
          "res": rng.uniform(0.2, 1, (B, N)), "std_u": rng.uniform(0, 0.5, N), "spikes": np.zeros((B, N)),
          "out": np.zeros((B, N)), "rate": rng.uniform(0, 50, (B, N)), "spike_counts": np.zeros((B, N))}
    st = {k: np.ascontiguousarray(v, np.float32) for k, v in st.items()}
    st["v"][0, :4] = -45.0
    ts = {k: torch.from_numpy(v.copy()) for k, v in st.items()}
    for it in range(20):
        rnd = rng.random((B, N)).astype(np.float32)
        oracle.lif(st, flags, P, rnd)
        torch_reference_step(ts, P, torch.from_numpy(rnd), flags, dt)
        for k in ("v", "refrac", "spikes", "out", "rate", "spike_counts", "adapt", "res"):
            assert torch.equal(torch.from_numpy(st[k]), ts[k]), (k, it)


def run_lif(st, steps, P, flags=0):
    for _ in range(steps):
        oracle.lif(st, flags, P)


def single(v=-52.0, g=0.0, drive=0.0, refrac=0.0):
    f = lambda x: np.full((1, 1), x, np.float32)
    return {"v": f(v), "g": f(g), "drive": f(drive), "adapt": f(0), "refrac": f(refrac), "active": np.ones(1, np.float32),
            "poisson_p": f(0), "res": f(1), "std_u": np.zeros(1, np.float32), "spikes": f(0), "out": f(0),
            "rate": f(0), "spike_counts": f(0)}


def test_lif_exponential_euler_closed_form():
    """Held input I below threshold: v_k = (v_rest + I) + (v0 - v_rest - I) a_m^k."""
    P = P_vec()
    st = single(v=-52.0, g=4.0)
    run_lif(st, 300, P)
    a_m = float(P[3])
    expect = (-52.0 + 4.0) + (-52.0 - (-52.0 + 4.0)) * a_m**300
    assert abs(float(st["v"][0, 0]) - expect) < 2e-4
    assert st["spike_counts"].sum() == 0


def test_lif_threshold_equality_fires_and_resets():
    P = P_vec()
    st = single(v=-45.0, g=7.0)          # target = -45 exactly, v stays at -45: v >= v_th fires
    oracle.lif(st, 4, P)
    assert st["spikes"][0, 0] == 1.0 and st["v"][0, 0] == -52.0 and st["refrac"][0, 0] == np.float32(2.2)


def test_refractory_hold_is_five_steps_at_dt_half():
    """The reference's documented known answer: t_ref 2.2 ms at dt 0.5 holds a cell 5 steps and it fires again on
    the 6th (docs/NOTES.md:60-62)."""
    P = P_vec(dt=0.5)
    st = single(v=-40.0, g=1.0e4)        # drives back over threshold on the first step after the hold
    oracle.lif(st, 4, P)
    held = 0
    while True:
        oracle.lif(st, 4, P)
        if st["spikes"][0, 0]:
            break
        held += 1
    assert held == 5


def test_frozen_cells_only_fire_when_forced():
    P = P_vec()
    st = single(v=-30.0)
    st["active"][:] = 0
    oracle.lif(st, 4, P)
    assert st["spikes"][0, 0] == 0
    st["poisson_p"][:] = 0.5
    oracle.lif(st, 1 | 4, P, rnd=np.full((1, 1), 0.25, np.float32))
    assert st["spikes"][0, 0] == 1


def test_compact_skip_leaves_resting_frozen_cells_untouched():
    P = P_vec()
    st = single()
    st["active"][:] = 0
    st["rate"][:] = 0
    before = {k: v.copy() for k, v in st.items()}
    oracle.lif(st, 32, P)
    for k in st:
        assert np.array_equal(st[k], before[k]), k


# ------------------------------------------------------------------ synaptic input
def test_events_match_index_add_order_and_float64():
    rng = np.random.default_rng(0)
    n, B = 300, 2
    W = random_csr(rng, n, n, 0.1) * 7
    Wc = sp.csc_matrix(W)
    Wc.sort_indices()
    x = (rng.random((B, n)) < 0.2).astype(np.float32)
    x[:, ::5] *= 0.37
    g = rng.normal(0, 2, (B, n)).astype(np.float32)
    ref = g.copy()
    oracle.events(Wc.indptr, Wc.indices, Wc.data, x, ref)
    # the reference CPU path: nonzero(x) in (b, pre) order, out-edges in CSC order, serial index_add_
    t = torch.from_numpy(g.copy()).view(-1)
    nz = torch.nonzero(torch.from_numpy(x))
    ptr, post, w = (torch.from_numpy(a) for a in (Wc.indptr.astype(np.int64), Wc.indices.astype(np.int64), Wc.data))
    xt = torch.from_numpy(x)
    for b, pre in nz.tolist():
        e = torch.arange(int(ptr[pre]), int(ptr[pre + 1]))
        t.index_add_(0, post[e] + b * n, w[e] * xt[b, pre])
    assert torch.equal(t.view(B, n), torch.from_numpy(ref))
    exact = g.astype(np.float64) + (W.astype(np.float64) @ x.astype(np.float64).T).T
    np.testing.assert_allclose(ref, exact, rtol=1e-5, atol=1e-4)


# ------------------------------------------------------------------ CSR products
def test_spmv_close_to_float64_and_chunk_order():
    rng = np.random.default_rng(1)
    M = random_csr(rng, 500, 400, 0.05, long_rows=(3,), long_len=390)
    x = rng.standard_normal((2, 400)).astype(np.float32)
    y = oracle.spmv(M.indptr, M.indices, M.data, x)
    exact = (M.astype(np.float64) @ x.astype(np.float64).T).T
    np.testing.assert_allclose(y, exact, rtol=1e-5, atol=1e-5)
    # chunk structure: the long row's value is the fold of its 16-entry chunk sums, not the plain sequential sum
    r = 3
    s, e = M.indptr[r], M.indptr[r + 1]
    prods = (M.data[s:e] * x[0, M.indices[s:e]]).astype(np.float32)
    total = np.float32(0)
    for c0 in range(0, e - s, 16):
        part = np.float32(0)
        for q in prods[c0:c0 + 16]:
            part = np.float32(part + q)
        total = np.float32(total + part)
    assert y[0, r] == total


def test_half_rounding_matches_numpy_float16():
    rng = np.random.default_rng(2)
    vals = np.concatenate([rng.standard_normal(5000) * 10 ** rng.uniform(-8, 4.5, 5000),
                           [0.0, -0.0, 65504.0, 65519.0, 1e-8, 6e-8, 2.98e-8, 1.0 + 2**-11, 1.0 + 3 * 2**-11]])
    x = vals.astype(np.float32)[None]
    n = x.shape[1]
    I = sp.identity(n, dtype=np.float32, format="csr")
    y = oracle.spmv(I.indptr, I.indices, I.data, x, half=True)
    assert np.array_equal(y[0], x[0].astype(np.float16).astype(np.float32))


# ------------------------------------------------------------------ optic lobe
def test_optic_substep_matches_reference_torch_ops():
    """The expressions of flyverse/optic.py:528-551 (and the test_cuda.py:42-63 formula) in torch float32."""
    rng = np.random.default_rng(4)
    B, n = 3, 259
    v = (rng.random((B, n)) - .5).astype(np.float32)
    adapt = rng.random((B, n)).astype(np.float32)
    a = rng.random(n).astype(np.float32)
    b = rng.random(n).astype(np.float32)
    tv, ta, tb, taa = (torch.from_numpy(q.copy()) for q in (v, adapt, b, a))
    dr = oracle.optic_dr(v, b)
    for _ in range(10):
        y, pr, sk = (rng.random((B, n)).astype(np.float32) for _ in range(3))
        oracle.optic_update(v, adapt, dr, y, pr, sk, a, b, 1.0, 1.0, 0.997)
        tdr = (tv + tb[None]).clamp(0.0, 1.0) - tb[None]
        inp = 1.0 * torch.from_numpy(y) + torch.from_numpy(pr) - 1.0 * ta
        inp = inp + torch.from_numpy(sk)
        torch.add(inp, (tv - inp) * taa[None], out=tv)
        torch.add(tdr, (ta - tdr) * 0.997, out=ta)
    assert torch.equal(torch.from_numpy(v), tv)
    assert torch.equal(torch.from_numpy(adapt), ta)
    assert torch.equal(torch.from_numpy(dr), (tv + tb[None]).clamp(0.0, 1.0) - tb[None])


def test_photoreceptor_matches_reference_torch_ops():
    """flyverse/optic.py:503-513 with the identity averaging matrix, torch float32."""
    rng = np.random.default_rng(5)
    B, n = 2, 300
    sens = rng.random((n, 4)).astype(np.float32)
    I_lp, I_mean = np.zeros((B, 5 * n), np.float32), np.zeros((B, 5 * n), np.float32)
    fresh = np.ones(B, np.uint8)
    tl, tm = torch.zeros(B, n), torch.zeros(B, n)
    tfresh = torch.ones(B, dtype=torch.bool)
    frame_ms = 10.0
    a_lp = float(np.exp(-frame_ms / 10.0))
    a_ad = float(np.exp(-frame_ms / 300.0))
    coefs = np.array([a_lp, 1 - a_lp, a_ad, 1 - a_ad, 0.02, 2.0], np.float32)
    for _ in range(5):
        rad = rng.random((B, n, 4)).astype(np.float32)
        got = oracle.photoreceptor(rad, sens, I_lp, I_mean, fresh, coefs)
        r = torch.from_numpy(rad)
        I = r[:, :, 0] * torch.from_numpy(sens)[None, :, 0]
        for c in range(1, 4):               # channel order (the oracle's defined order of the 4-term sum)
            I = I + r[:, :, c] * torch.from_numpy(sens)[None, :, c]
        torch.where(tfresh[:, None], I, tl, out=tl)
        torch.where(tfresh[:, None], I, tm, out=tm)
        tfresh[:] = False
        torch.add(a_lp * tl, (1 - a_lp) * I, out=tl)
        torch.add(a_ad * tm, (1 - a_ad) * tl, out=tm)
        contrast = ((tl - tm) / (tm + 0.02)).clamp(-1.0, 2.0)
        assert torch.equal(torch.from_numpy(got), contrast)
    assert fresh.sum() == 0


def _reference_photoreceptor_torch(retina, rad, tl, tm, tfresh, coefs):
    """flyverse/optic.py:496-517 written with torch float32 ops on the reference's own averaging matrix
    (optic.py:315-325); the channel sum in channel order."""
    pr_col, fam = retina.pr_column, retina.pr_family
    n_col, n_pr = retina.n_col, len(pr_col)
    cnt = np.zeros((n_col, 5), np.float32)
    np.add.at(cnt, (pr_col, fam), 1.0)
    rows = pr_col * 5 + fam
    A = sp.csr_matrix((1.0 / cnt[pr_col, fam], (rows, np.arange(n_pr))), shape=(n_col * 5, n_pr))
    At = torch.sparse_csr_tensor(torch.from_numpy(A.indptr.astype(np.int64)), torch.from_numpy(A.indices.astype(np.int64)),
                                 torch.from_numpy(A.data.astype(np.float32)), size=A.shape)
    has = np.zeros((n_col, 5), bool)
    np.add.at(has, (pr_col, fam), True)
    r = torch.from_numpy(rad)[:, torch.from_numpy(pr_col)]
    sens = torch.from_numpy(retina.pr_sens)
    I_pr = r[:, :, 0] * sens[None, :, 0]
    for c in range(1, 4):
        I_pr = I_pr + r[:, :, c] * sens[None, :, c]
    # the averaging product in the order every CSR product of this build defines (its rows here are shorter than
    # one 16-entry chunk: a sequential fold from +0, each product rounded); torch's own sparse product on the CPU
    # (MKL, fused multiply-adds) is the reference's arithmetic and agrees to rounding
    I = torch.zeros(I_pr.shape[0], A.shape[0])
    for k in range(A.shape[0]):
        for e in range(A.indptr[k], A.indptr[k + 1]):
            I[:, k] = I[:, k] + torch.tensor(A.data[e], dtype=torch.float32) * I_pr[:, A.indices[e]]
    torch.testing.assert_close((At @ I_pr.T.contiguous()).T, I, rtol=1e-6, atol=0)
    torch.where(tfresh[:, None], I, tl, out=tl)
    torch.where(tfresh[:, None], I, tm, out=tm)
    tfresh[:] = False
    k = [float(x) for x in coefs]                   # python floats: torch rounds them to float32, as coefs holds
    torch.add(k[0] * tl, k[1] * I, out=tl)
    torch.add(k[2] * tm, k[3] * tl, out=tm)
    contrast = ((tl - tm) / (tm + float(coefs[4]))).clamp(-1.0, float(coefs[5]))
    contrast = torch.where(torch.from_numpy(has.reshape(-1))[None], contrast, torch.zeros_like(contrast))
    return contrast, contrast[:, torch.from_numpy(rows)], A, has.reshape(-1)


def test_photoreceptor_cells_match_reference_torch_ops_and_float64():
    """Several photoreceptors per (column, family) cell, empty cells and columns without photoreceptors: the oracle
    against flyverse/optic.py:496-517 in torch float32 with the reference's averaging matrix (bit for bit), and
    against a float64 numpy restatement within 1e-5 relative."""
    rng = np.random.default_rng(11)
    B, n_col, n_pr = 3, 40, 150
    pr_col = rng.integers(0, n_col - 3, n_pr)                 # the last three columns have no photoreceptor
    fam = rng.choice(5, n_pr, p=[0.6, 0.1, 0.1, 0.1, 0.1])
    retina = Retina(np.arange(n_pr), rng.random((n_pr, 4)).astype(np.float32), pr_col, fam, n_col)
    cells = retina.cells()
    assert np.bincount(retina.pr_column * 5 + retina.pr_family).max() >= 5
    I_lp, I_mean = np.zeros((B, 5 * n_col), np.float32), np.zeros((B, 5 * n_col), np.float32)
    fresh = np.ones(B, np.uint8)
    tl, tm = torch.zeros(B, 5 * n_col), torch.zeros(B, 5 * n_col)
    tfresh = torch.ones(B, dtype=torch.bool)
    frame_ms = 10.0
    a_lp, a_ad = float(np.exp(-frame_ms / 10.0)), float(np.exp(-frame_ms / 300.0))
    coefs = np.array([a_lp, 1 - a_lp, a_ad, 1 - a_ad, 0.02, 2.0], np.float32)
    d_lp = d_mean = None
    for frame in range(6):
        rad = rng.random((B, n_col, 4)).astype(np.float32)
        con = np.empty((B, 5 * n_col), np.float32)
        got = oracle.photoreceptor(rad, retina.pr_sens, I_lp, I_mean, fresh, coefs, cells, contrast=con)
        contrast, a_pr, A, has = _reference_photoreceptor_torch(retina, rad, tl, tm, tfresh, coefs)
        assert torch.equal(torch.from_numpy(con), contrast) and torch.equal(torch.from_numpy(got), a_pr)
        assert torch.equal(torch.from_numpy(I_lp), tl) and torch.equal(torch.from_numpy(I_mean), tm)
        # float64, the same formulas
        Ip = (rad.astype(np.float64)[:, retina.pr_column] * retina.pr_sens.astype(np.float64)).sum(-1)
        I = (A.astype(np.float64) @ Ip.T).T
        if frame == 0:
            d_lp, d_mean = I.copy(), I.copy()
        d_lp = a_lp * d_lp + (1 - a_lp) * I
        d_mean = a_ad * d_mean + (1 - a_ad) * d_lp
        c64 = np.where(has[None], np.clip((d_lp - d_mean) / (d_mean + 0.02), -1.0, 2.0), 0.0)
        np.testing.assert_allclose(I_lp, d_lp, rtol=1e-6, atol=1e-7)
        np.testing.assert_allclose(con, c64, rtol=1e-5, atol=1e-6)
        assert (con[:, ~has] == 0).all() and np.array_equal(got, con[:, retina.pr_column * 5 + retina.pr_family])


def test_feedback_and_drive_out_and_means():
    rng = np.random.default_rng(6)
    rate = rng.uniform(-10, 400, (2, 50)).astype(np.float32)
    idx = np.array([1, 4, 9, 30], np.int32)
    s = oracle.feedback(rate, idx)
    assert torch.equal(torch.from_numpy(s), (torch.from_numpy(rate)[:, idx] / 100.0).clamp(0, 3))
    y = rng.normal(0, 1, (2, 4)).astype(np.float32)
    d = oracle.drive_out(y, idx, 50, 100.0, 35.0)
    expect = torch.zeros(2, 50)
    expect[:, idx] = (100.0 * torch.from_numpy(y)).clamp(-35, 35)
    assert torch.equal(torch.from_numpy(d), expect)
    ptr = np.array([0, 2, 2, 4], np.int32)
    out = oracle.group_means(ptr, idx, rate)
    assert np.array_equal(out[:, 0], rate[:, idx[:2]].astype(np.float64).sum(1) / 2) and (out[:, 1] == 0).all()
