"""The optic-lobe adjoint.  CPU: the float64 torch restatement (oracle/optic_grad.py) is pinned by central finite
differences.  GPU: the FP32 forward + hand-written adjoint (flyhip.adjoint) against float64 autograd of that
restatement on the same inputs: loss within 1e-5 relative, every gradient group within 1e-4 relative (norm of the
difference over the norm of the float64 gradient); checkpoint spacing (or keeping the whole trajectory) does not
change a bit, nor does a repeated backward; the autograd wrapper returns the same gradients."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle.optic_grad import loss64
from flyhip import optic_data


def small_problem(n=300, K=5, T=24, nnz=6000, seed=0, B=32):
    rng = np.random.default_rng(seed)
    W0 = optic_data._normalised(rng, n, n, nnz)
    ut = rng.integers(0, K, n)
    P = (rng.normal(0, 0.3, (T, n, B)) + 0.2 * np.sin(np.arange(T))[:, None, None]).astype(np.float32)
    readout = np.sort(rng.choice(n, max(4, n // 10), replace=False))
    tau = (20.0 * np.exp(rng.normal(0, 0.3, K))).astype(np.float32)
    bias = rng.normal(0, 0.1, K).astype(np.float32)
    gain = np.exp(rng.normal(0, 0.2, (K, K))).astype(np.float32)
    target = rng.uniform(0, 1, (T, len(readout), B)).astype(np.float32)
    return W0, ut, P, target, readout, tau, bias, gain


def grads64(W0, ut, P, target, readout, tau, bias, gain):
    th = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in (tau, bias, gain)]
    loss = loss64(*th, W0, ut, P, target, readout)
    loss.backward()
    return loss.item(), [x.grad.numpy() for x in th]


def test_float64_reference_matches_finite_differences():
    W0, ut, P, target, readout, tau, bias, gain = small_problem(n=40, K=3, T=12, nnz=300, B=4)
    P = P[:, :, :4].copy()
    target = target[:, :, :4].copy()
    _, (gt, gb, gg) = grads64(W0, ut, P, target, readout, tau, bias, gain)
    h = 1e-6
    params = [np.asarray(tau, np.float64), np.asarray(bias, np.float64), np.asarray(gain, np.float64)]

    def L(ps):
        return float(loss64(*[torch.tensor(p) for p in ps], W0, ut, P, target, readout))

    for which, g in zip(range(3), (gt, gb, gg)):
        flat = params[which].reshape(-1)
        for k in range(flat.size):
            plus = [p.copy() for p in params]
            minus = [p.copy() for p in params]
            plus[which].reshape(-1)[k] += h
            minus[which].reshape(-1)[k] -= h
            fd = (L(plus) - L(minus)) / (2 * h)
            assert abs(fd - g.reshape(-1)[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (which, k, fd, g.reshape(-1)[k])


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
@pytest.mark.parametrize("checkpoint", [1, 7, 24, None])
def test_adjoint_matches_float64_autograd(checkpoint):
    from flyhip.adjoint import OpticAdjoint
    W0, ut, P, target, readout, tau, bias, gain = small_problem()
    L64, (gt, gb, gg) = grads64(W0, ut, P, target, readout, tau, bias, gain)
    prob = OpticAdjoint(W0, ut, 5, P, target, readout, checkpoint=checkpoint, device="cuda")
    t = lambda x: torch.as_tensor(x, device="cuda")
    L = float(prob.loss(t(tau), t(bias), t(gain)))
    g = [x.cpu().numpy() for x in prob.gradients()]
    assert abs(L - L64) <= 1e-5 * abs(L64)
    for name, got, ref in (("tau", g[0], gt), ("bias", g[1], gb), ("gain", g[2], gg)):
        assert rel(got, ref) <= 1e-4, (name, rel(got, ref))


@pytest.mark.gpu
def test_adjoint_is_deterministic_and_checkpoint_invariant():
    from flyhip.adjoint import OpticAdjoint, OpticLossFn
    W0, ut, P, target, readout, tau, bias, gain = small_problem(seed=3)
    t = lambda x: torch.as_tensor(x, device="cuda")
    results = []
    for ck in (5, 5, 24, None, 7):
        prob = OpticAdjoint(W0, ut, 5, P, target, readout, checkpoint=ck, device="cuda")
        L = prob.loss(t(tau), t(bias), t(gain)).clone()
        results.append((L, [x.clone() for x in prob.gradients()]))
        # a second backward recomputes the segments the first one took from the forward: the same bits
        results.append((L, [x.clone() for x in prob.gradients()]))
    for L, g in results[1:]:
        assert torch.equal(L, results[0][0])
        for a, b in zip(g, results[0][1]):
            assert torch.equal(a, b)
    # through torch autograd
    th = [t(x).clone().requires_grad_(True) for x in (tau, bias, gain)]
    loss = OpticLossFn.apply(*th, prob)
    loss.backward()
    for a, b in zip(th, results[0][1]):
        assert torch.equal(a.grad, b)


def _half_sums16(p):
    """The per-entry sums over 32 lanes of the adjoint's transposing butterfly, on the CPU: p (..., 32 lanes, 16
    entries) float32 -> (..., 16), each entry the value lane 2m ends with (float32 adds in the kernel's tree)."""
    lanes = np.arange(32)
    vals = p.astype(np.float32)
    for bit, half in ((16, 8), (8, 4), (4, 2), (2, 1)):
        up = (lanes & bit) != 0
        keep = np.where(up[:, None], vals[..., half:2 * half], vals[..., :half])
        give = np.where(up[:, None], vals[..., :half], vals[..., half:2 * half])
        vals = (keep + give[..., lanes ^ bit, :]).astype(np.float32)
    v = vals[..., 0]
    v = (v + v[..., lanes ^ 1]).astype(np.float32)
    return v[..., 0::2]


@pytest.mark.gpu
def test_adjoint_edge_partials_from_the_transposed_product():
    """The per-edge gain partials that launch_adj_back_rows adds from its own gathers: for each entry (W row r) of
    unit j's column, eg += sum over the 32 movies of ly[r, b] * dr[j, b] in the kernel's fixed butterfly order (bit
    for bit against a CPU emulation), within 1e-5 of float64; pads and a null eg are left alone."""
    import ctypes as ct
    from flyhip.adjoint import OpticAdjoint
    W0, ut, P, target, readout, tau, bias, gain = small_problem(n=200, nnz=9000, seed=4)
    prob = OpticAdjoint(W0, ut, 5, P, target, readout, device="cuda")
    t = lambda x: torch.as_tensor(x, device="cuda")
    prob.loss(t(tau), t(bias), t(gain))
    rng = np.random.default_rng(2)
    ly = rng.normal(0, 1, (prob.n, 32)).astype(np.float32)
    dr = rng.normal(0, 1, (prob.n, 32)).astype(np.float32)
    prob.ly.copy_(t(ly))
    prob.S_dr[0].copy_(t(dr))
    eg0 = rng.normal(0, 1, prob.eg.numel()).astype(np.float32)
    eg = t(eg0.copy())
    f_back = (ct.c_float * 4)(1.0, prob.a_ad, 0.01, 1.0)
    prob._back_rows(0, 1, f_back, eg)
    torch.cuda.synchronize()
    got = eg.cpu().numpy()
    bptr, bidx = prob.bptr.cpu().numpy(), prob.bidx.cpu().numpy()
    unit = np.repeat(np.arange(prob.n), np.diff(bptr))                 # the W^T row (unit) of each chunk
    chunks = bidx.reshape(-1, 16)
    prods = (ly[np.maximum(chunks, 0)] * dr[unit][:, None, :]).astype(np.float32)    # (chunks, 16, 32)
    prods[chunks < 0] = 0.0
    sums = _half_sums16(np.transpose(prods, (0, 2, 1)))                 # (chunks, 16)
    want = eg0.reshape(-1, 16).copy()
    real = chunks >= 0
    want[real] = (want[real] + sums[real]).astype(np.float32)
    assert np.array_equal(got, want.reshape(-1))
    exact = (ly[np.maximum(chunks, 0)].astype(np.float64) * dr[unit][:, None, :]).sum(-1)
    np.testing.assert_allclose(sums[real], exact[real], rtol=1e-5, atol=1e-5)
    before = prob.lv.clone()
    prob._back_rows(0, 1, f_back, None)                                 # no eg: the step's other outputs only
    assert not torch.equal(before, prob.lv)


@pytest.mark.gpu
def test_adjoint_forward_is_the_inference_substep_bit_for_bit():
    """The training forward evaluates the inference expressions (the invariant of the reference's
    tests/test_surrogate.py:178-183, grad path == inference path bit for bit): every step's v, adapt and dr of
    OpticAdjoint's FP32 forward equal the CPU oracle's inference substep (oracle.spmv in the chunked order of the
    CSR products, then oracle.optic_update, flyverse/optic.py:525-551 with no spiking feedback), given the same
    per-unit coefficients and shaped weights."""
    import oracle
    from flyhip.adjoint import OpticAdjoint
    n, K, T = 3000, 7, 12
    W0, ut, P, target, readout, tau, bias, gain = small_problem(n=n, K=K, T=T, nnz=90_000, seed=5)
    prob = OpticAdjoint(W0, ut, K, P, target, readout, checkpoint=None, device="cuda")
    t = lambda x: torch.as_tensor(x, device="cuda")
    prob.loss(t(tau), t(bias), t(gain))
    torch.cuda.synchronize()
    a, b = prob.a.cpu().numpy(), prob.b.cpu().numpy()
    fidx, fw = prob.fidx.cpu().numpy(), prob.fw.cpu().numpy()
    w = fw[fidx >= 0]                                          # the shaped weights in CSR order
    W = sp.csr_matrix(W0, dtype=np.float32)
    W.sum_duplicates()
    W.sort_indices()
    assert np.array_equal(fidx[fidx >= 0], W.indices)
    ptr = W.indptr.astype(np.int64)
    v = np.zeros((32, n), np.float32)
    adapt = np.zeros((32, n), np.float32)
    dr = oracle.optic_dr(v, b)
    zero = np.zeros((32, n), np.float32)
    c = prob.c
    for step in range(T):
        y = oracle.spmv(ptr, W.indices, w, dr)
        oracle.optic_update(v, adapt, dr, y, np.ascontiguousarray(P[step].T), zero, a, b, c.gain_rr, c.adapt_gain,
                            prob.a_ad)
        for name, got, want in (("v", prob.S_v[step + 1], v), ("adapt", prob.S_a[step + 1], adapt),
                                ("dr", prob.S_dr[step + 1], dr)):
            assert np.array_equal(got.cpu().numpy().T, want), (step, name)


@pytest.mark.gpu
def test_full_size_adjoint_matches_the_float64_fixture():
    """The full-size training problem (89,390 units, 5,995,830 edges, 65 types, 32 movies, 200 steps, checkpoints
    every 20 steps) against the committed float64 autograd loss and gradient (tests/golden/make_c4_gradient.py):
    loss within 1e-5 relative, each gradient group within 1e-4 relative; keeping the whole trajectory instead gives
    the same bits."""
    from pathlib import Path
    from flyhip.adjoint import OpticAdjoint
    from helpers import c4_problem
    fx = np.load(Path(__file__).parent / "golden" / "c4_gradient.npz")
    W0, ut, P, target, readout, tau, bias, gain = c4_problem()
    t = lambda x: torch.as_tensor(x, device="cuda")
    results = []
    for ck in (20, None):
        prob = OpticAdjoint(W0, ut, 65, P, target, readout, checkpoint=ck, device="cuda")
        L = prob.loss(t(tau), t(bias), t(gain)).clone()
        g = [x.clone() for x in prob.gradients()]
        results.append((L, g))
        del prob
        torch.cuda.empty_cache()
    (L, g), (L2, g2) = results
    assert torch.equal(L, L2)
    for a_, b_ in zip(g, g2):
        assert torch.equal(a_, b_)
    L64 = float(fx["loss"])
    assert abs(float(L) - L64) <= 1e-5 * abs(L64), (float(L), L64)
    for name, got, ref in (("tau", g[0], fx["g_tau"]), ("bias", g[1], fx["g_bias"]), ("gain", g[2], fx["g_gain"])):
        assert rel(got.cpu().numpy(), ref) <= 1e-4, (name, rel(got.cpu().numpy(), ref))


def test_c4_fixture_gradient_agrees_with_its_finite_differences():
    """The committed full-size float64 gradient against the float64 central differences of the same loss that the
    fixture script took for two entries of each parameter group: within 1e-4 of the entry (floor 1e-6 of the group's
    largest entry).  Not tighter: a bias shifts units across the kinks of clip(v + b, 0, 1), where a difference
    quotient averages the slopes on both sides while autograd takes torch's one-sided clamp gradient."""
    from pathlib import Path
    fx = np.load(Path(__file__).parent / "golden" / "c4_gradient.npz")
    groups = (fx["g_tau"].reshape(-1), fx["g_bias"].reshape(-1), fx["g_gain"].reshape(-1))
    assert fx["g_gain"].shape == (65, 65) and groups[0].size == 65 and groups[1].size == 65
    assert sorted(set(int(g) for g in fx["fd_group"])) == [0, 1, 2]
    for grp, k, fd in zip(fx["fd_group"], fx["fd_index"], fx["fd_value"]):
        g = groups[int(grp)]
        tol = 1e-4 * max(abs(g[int(k)]), 1e-2 * np.abs(g).max())
        assert abs(fd - g[int(k)]) <= tol, (int(grp), int(k), fd, g[int(k)])
