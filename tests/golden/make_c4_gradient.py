"""Writes tests/golden/c4_gradient.npz: the float64 loss and gradient (65 tau + 65 bias + 65 x 65 gains) of the
full-size differentiable optic-lobe problem (tests/helpers.c4_problem: 89,390 units, 5,995,830 edges, 65 types,
B = 32 movies, T = 200 steps at dt 5 ms, 734 readout units), from float64 torch autograd of oracle.optic_grad.loss64
(one checkpoint per step), plus float64 central finite differences of the same loss for a few entries of each
parameter group.  Runs on the CPU (about 40 minutes on 8 cores); the GPU test compares the FP32 adjoint with it.

    python tests/golden/make_c4_gradient.py
"""
import sys
import time
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
for p in (ROOT, ROOT / "flyverse-core_amd", ROOT / "tests"):
    sys.path.insert(0, str(p))

from helpers import c4_problem  # noqa: E402
from oracle.optic_grad import loss64  # noqa: E402

# (group, flat index) of the finite-difference spot checks; group 0 tau, 1 bias, 2 gain
FD_ENTRIES = ((0, 3), (0, 41), (1, 7), (1, 60), (2, 12 * 65 + 30), (2, 50 * 65 + 50))
FD_STEP = (1e-4, 1e-6, 1e-6)      # tau is in ms (about 20), bias and gain are O(0.1) and O(1)


def main():
    torch.set_num_threads(max(1, torch.get_num_threads()))
    t0 = time.time()
    W0, ut, P, target, readout, tau, bias, gain = c4_problem()
    print(f"problem built in {time.time() - t0:.0f} s: n {W0.shape[0]}, nnz {W0.nnz}, P {P.shape}", flush=True)
    args = (W0, ut, P, target, readout)
    th = [torch.tensor(np.asarray(x, np.float64), requires_grad=True) for x in (tau, bias, gain)]
    t0 = time.time()
    loss = loss64(*th, *args, per_step_checkpoint=True)
    print(f"float64 forward {time.time() - t0:.0f} s, loss {loss.item():.17g}", flush=True)
    t0 = time.time()
    loss.backward()
    print(f"float64 backward {time.time() - t0:.0f} s", flush=True)
    grads = [x.grad.numpy().copy() for x in th]
    base = [np.asarray(x, np.float64) for x in (tau, bias, gain)]
    fd = []
    with torch.no_grad():
        for group, k in FD_ENTRIES:
            h = FD_STEP[group]
            vals = []
            for sgn in (1.0, -1.0):
                ps = [b.copy() for b in base]
                ps[group].reshape(-1)[k] += sgn * h
                vals.append(float(loss64(*[torch.tensor(p) for p in ps], *args)))
            fd.append((vals[0] - vals[1]) / (2 * h))
            print(f"FD group {group} entry {k}: {fd[-1]:.12g} vs autograd {grads[group].reshape(-1)[k]:.12g}",
                  flush=True)
    np.savez(HERE / "c4_gradient.npz", loss=np.float64(loss.item()), g_tau=grads[0], g_bias=grads[1],
             g_gain=grads[2], fd_group=np.array([g for g, _ in FD_ENTRIES]), fd_index=np.array([k for _, k in FD_ENTRIES]),
             fd_step=np.array([FD_STEP[g] for g, _ in FD_ENTRIES]), fd_value=np.array(fd))
    print("wrote", HERE / "c4_gradient.npz")


if __name__ == "__main__":
    main()
