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


@pytest.mark.gpu
@pytest.mark.parametrize("nnz", [1, 63, 64, 200, 4097])
def test_adjoint_edge_products_bit_exact(nnz):
    """launch_adj_back_edges: eg[e] += sum over the 32 movies, in movie order, of ly[row_e, b] * dr[col_e, b] (each
    product rounded), for edge counts around the 64-edge wavefront tiles; torch float32 in the same order."""
    import ctypes as ct
    from flyhip.adjoint import _lib
    rng = np.random.default_rng(nnz)
    n = 97
    row_of = np.sort(rng.integers(0, n, nnz)).astype(np.int32)
    idx = rng.integers(0, n, nnz).astype(np.int32)
    ly = rng.normal(0, 1, (n, 32)).astype(np.float32)
    dr = rng.normal(0, 1, (n, 32)).astype(np.float32)
    eg0 = rng.normal(0, 1, nnz).astype(np.float32)
    t = [torch.as_tensor(x, device="cuda") for x in (row_of, idx, ly, dr, eg0)]
    arr = (ct.c_void_p * 5)(*(x.data_ptr() for x in t))
    assert _lib(t[0].device).launch_adj_back_edges(arr, nnz, torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    L, D = torch.from_numpy(ly)[torch.from_numpy(row_of).long()], torch.from_numpy(dr)[torch.from_numpy(idx).long()]
    s = torch.zeros(nnz)
    for b in range(32):
        s = s + L[:, b] * D[:, b]
    assert torch.equal(t[4].cpu(), torch.from_numpy(eg0) + s)
