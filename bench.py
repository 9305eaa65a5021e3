"""Brain-step benchmark: batched sim-s/wall-s at 167k neurons (B = 64) against the plain single-fly path (B = 1).

Workload (synthetic, seeded): the full-scale fly graph (167,106 neurons, ~25.4 M edges, flyhip.synth.fly_graph),
Shiu LIF at dt 0.1 ms coupled to the 89,390-unit optic lobe every 10 ms frame (10 optic substeps of 1 ms).
  * B = 64: flyhip.batch.BatchBrain, 64 lanes fed from a queue of rollouts, each with its own Poisson seed,
    photoreceptor movie seed, 300 Poisson inputs at U[0, 200] Hz and a length uniform over 2,000-10,000 LIF steps
    (whole 10 ms frames); finished lanes are refilled.  One bench step = one 10 ms frame of all 64 lanes, replayed
    from a captured HIP graph.  value = simulated fly-seconds / wall-seconds over the timed frames.
  * B = 1: flyhip.FlyBrain(batch=1, cuda_graphs=True), the plain kernels (window launches of up to 18 LIF steps,
    launch_lif_window; the fused frame-start / frame-end optic launches around 10 substeps), one captured graph per
    frame shape; reported as b1_sim_s_per_wall_s and the ratio.
  * roofline: the batched LIF launch (csrc/batch.hip k_batch_lif, <= n_delay steps per launch): one frame of its
    launches timed with HIP events on the launch stream right after the timed region; algorithmic bytes per launch
    from the per-step LIF and synaptic-input byte counts (lif_block_bytes).  The fused optic substep of the same
    engine (k_optic_substep_fm) is timed the same way and reported as optic_substep_b64_us / _frac.
  * packed: 128 LIF subgraphs (5k-80k cells, log-uniform; in-degree of the full graph) packed into one CSR and
    stepped 5,000 steps by one replay of a captured graph of window launches (csrc/packed.hip), window and grid
    from flyhip.autotune.
  * cpu_baseline: the CPU path on all the host's cores, 4 flies of the same workload for a bounded sample of frames:
    the C oracle with OpenMP, and the reference's step in torch CPU ops (oracle.torch_ref); the faster is the value.
  * --only abi (not part of the default run): the reference's six-entry C ABI at full scale, B = 1 and 8.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "flyverse-core_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_rollouts(layout, n, seed, frame_steps=100, lo=2000, hi=10000):
    from flyhip import synth
    from flyhip.batch import Rollout
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        idx, hz = synth.fly_inputs(int(rng.integers(0, 2**31)), layout, 300, 200.0)
        frames = int(rng.integers(lo // frame_steps, hi // frame_steps + 1))
        out.append(Rollout(seed=int(rng.integers(1, 2**62)), movie_seed=int(rng.integers(1, 2**62)),
                           steps=frames * frame_steps, input_idx=idx, input_hz=hz))
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def lif_block_bytes(bb, K, B):
    """Algorithmic bytes of one k_batch_lif launch of K steps (a mean over launches may be fractional) for B live
    flies: per step, the LIF state traffic
    B*N_act*(32 + 8 + 1/8) and the synaptic input 8 * sum of out-degrees over the presynaptic cells that fired in
    any fly + B*N_act/8 for the spike bits (the per-step counts of the measurement plan).  The firing set is
    measured over the steps held in the ring."""
    import torch
    T = int(bb.engine_step.item())
    n = min(bb.ring_slots, T)
    slots = {t: bb.ring_slot(t) for t in range(T - n, T)}
    outdeg = np.bincount(bb.W_lif.indices, minlength=bb.n_rows).astype(np.float64)
    touched = []
    for t in range(T - n, T):
        fired = slots[t] != 0
        touched.append(outdeg[fired].sum())
    N = bb.n_rows
    per_step = B * N * (32 + 8 + 1 / 8) + 8 * float(np.mean(touched)) + B * N / 8
    act = float(np.mean([np.unpackbits(slots[t].view(np.uint8)).sum() for t in range(T - n, T)]))
    return K * per_step, act / (B * N)


def time_lif_launches(bb):
    """The k_batch_lif launches of one frame (every offset, as the captured frame issues them), each timed with HIP
    events on its launch stream.  Returns (mean duration in ms, mean steps per launch)."""
    import torch
    from flyhip.batch import _lib
    arr = bb.lif_pointers()
    stream = torch.cuda.current_stream()
    fn = _lib(bb.device).launch_batch_lif
    times, ks = [], []
    # the state just advances by one frame's LIF steps (the timed region is over)
    for off in range(0, bb.frame_steps, bb.steps_per_launch):
        K = min(bb.steps_per_launch, bb.frame_steps - off)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        err = fn(arr, bb._dims, off, K, stream.cuda_stream)
        e1.record(stream)
        if err:
            raise RuntimeError(f"launch_batch_lif: {err}")
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
        ks.append(K)
    return float(np.mean(times)), float(np.mean(ks))


def time_optic_substeps(bb, reps=20):
    """The fused fly-minor optic substep (csrc/batch.hip k_optic_substep_fm: W_rr product + update for 64 flies),
    `reps` launches timed with HIP events on the launch stream, over a copy of the engine's optic state (the engine's
    own state is left as it was).  Returns (mean us, algorithmic bytes 8 nnz + 4 (rows + 1) + 64 rows 4 B x 8)."""
    import ctypes as ct
    import torch
    from flyhip.batch import _lib
    M = bb.W_rr
    dr0, dr1, v, ad = (t.clone() for t in (bb.dr0, bb.dr1, bb.ov, bb.oadapt))
    ts = [M.ptr, M.idx, M.w, M.sched, dr0, dr1, v, ad, bb.pr_in, bb.spk_in, bb.o_a, bb.o_b, bb.live_mask]
    arr = (ct.c_void_p * len(ts))(*(t.data_ptr() for t in ts))
    fn = _lib(bb.device).launch_optic_substep_fm
    stream = torch.cuda.current_stream()
    args = (bb.n_rate, float(bb.op.gain_rr), float(bb.op.adapt_gain), float(bb.o_a_ad), bb.optic_grid(bb.n_rate),
            stream.cuda_stream)
    fn(arr, *args)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        err = fn(arr, *args)
        if err:
            raise RuntimeError(f"launch_optic_substep_fm: {err}")
    e1.record(stream)
    e1.synchronize()
    us = e0.elapsed_time(e1) * 1000 / reps
    nnz = int(M.nnz)
    return us, 8 * nnz + 4 * (bb.n_rate + 1) + 64 * bb.n_rate * 4 * 8


def spmv_roofline(conn, retina, op, reps=50):
    """The optic recurrent product y = W_rr x (flyverse/optic.py:474-483) at B=1 on the full-scale graph, grouped
    kernel (hip.CSR default): mean duration over `reps` launches timed with HIP events on the launch stream, and
    algorithmic bytes 8 nnz + 4 (rows + 1) + 4 (cols + rows)."""
    import torch
    from flyhip import hip
    from flyhip.optic import optic_matrices
    M = optic_matrices(conn, retina, op)["M_rr"]
    out = {}
    for kernel, rc in (("panel", 0), ("lds", 0), ("grouped", 0), ("chunked", 0)):
        csr = hip.CSR(M, "cuda", kernel=kernel, range_cols=rc)
        kernel = kernel + (f"{rc}" if rc else "")
        x = torch.rand(1, M.shape[1], device="cuda")
        y = csr.matvec(x)
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            csr.matvec(x, out=y)
        e1.record(stream)
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1000 / reps
        alg = 8 * M.nnz + 4 * (M.shape[0] + 1) + 4 * (M.shape[0] + M.shape[1])
        out[kernel] = (us, alg / (us * 1e-6) / 1e9)
        del csr
    best = min(out, key=lambda k: out[k][0])
    us, gbs = out[best]
    return {"spmv_wrr_b1_us": round(us, 2), "spmv_wrr_b1_gbs": round(gbs, 1),
            "spmv_wrr_b1_frac": round(gbs / HBM_PEAK_GBS, 4), "spmv_wrr_nnz": int(M.nnz), "spmv_wrr_kernel": best,
            **{f"spmv_wrr_b1_{k}_us": round(v[0], 2) for k, v in out.items()}}


def measured_traffic(shape):
    """HBM bytes per k_batch_lif launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary: the figure of
    launch shape `shape` (fold mode, steps per launch, workgroups per CU), else that of the recorded shape with the
    same fold mode and steps per launch and the nearest grid.  Returns (bytes, "fold,K,grid" it was measured at), or
    (None, None) when the summary holds no shape with that fold mode and step count."""
    path = ROOT / "profiles" / "hbm_traffic_k_batch_lif_b64.txt"
    if shape is None or not path.exists():
        return None, None
    rows = {}
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch":
            rows[tuple(int(x) for x in f[1].split(","))] = int(f[2])
    want = (shape["fold_mode"], shape["steps_per_launch"], shape["grid_per_cu"])
    same = [k for k in rows if k[:2] == want[:2]]
    if not same:
        return None, None
    key = min(same, key=lambda k: abs(k[2] - want[2]))
    return rows[key], ",".join(map(str, key))


def optic_training(steps=3, warmup=1):
    """The differentiable optic lobe at full scale (89,390 units, ~6 M edges, 65 types, 32 movies, 200 steps at dt
    5 ms, teacher targets on 734 readout units): forward + hand-written adjoint per training step, with checkpoints
    every 20 steps (the segments before the last are recomputed in the backward); the same step keeping the whole
    trajectory in HBM (nothing recomputed) is reported beside it."""
    import torch
    from flyhip import optic_data as od
    from flyhip.adjoint import OpticAdjoint
    W0, ut, W_rp, pos = od.optic_network()
    inten = od.grating_intensity(pos, od.T_STEPS)
    P = od.photoreceptor_drive(inten, W_rp)
    rng = np.random.default_rng(8)
    readout = np.sort(rng.choice(W0.shape[0], od.N_READOUT, replace=False))
    tau, bias, gain = od.initial_params()
    t = lambda x: torch.as_tensor(x, device="cuda")
    teacher = od.teacher_params(tau, bias, gain)
    th = [t(x) for x in (tau, bias, gain)]
    out = {}
    for ck in (None, 20):
        prob = OpticAdjoint(W0, ut, od.N_TYPES, P, np.zeros((od.T_STEPS, od.N_READOUT, 32), np.float32), readout,
                            checkpoint=ck)
        prob.target.copy_(prob.readout_trajectory(*[t(x) for x in teacher]))
        for _ in range(warmup):
            prob.loss(*th)
            prob.gradients()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            L = prob.loss(*th)
            g = prob.gradients()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        gn = float(torch.sqrt(sum((x.double() ** 2).sum() for x in g)))
        if ck is None:
            out["ol_fwd_bwd_ms_whole_trajectory"] = round(ms, 3)
        else:
            out.update({"ol_fwd_bwd_ms_per_step": round(ms, 3), "ol_us_per_simulated_step": round(ms * 1000 / od.T_STEPS, 2),
                        "ol_units": int(W0.shape[0]), "ol_edges": int(W0.nnz), "ol_movies": 32, "ol_steps": od.T_STEPS,
                        "ol_checkpoint": prob.checkpoint, "ol_loss": float(L), "ol_grad_norm": gn})
        del prob
        torch.cuda.empty_cache()
    return out


def packed_run(steps, log):
    """128 uneven subgraphs, each with its own seed and 300 Poisson inputs at U[0, 200] Hz, `steps` LIF steps in one
    launch after a tuning pass."""
    import torch
    from flyhip import synth
    from flyhip.autotune import tune_packed
    from flyhip.brain import LIFParams
    from flyhip.packed import PackedBrains
    t0 = time.perf_counter()
    mats, inputs, seeds = synth.packed_workload()
    lp = LIFParams.shiu(dt=0.1)
    pb = PackedBrains(mats, lp, seeds=seeds, inputs=inputs, device="cuda")
    del mats
    log(f"packed: {pb.k} subgraphs, {pb.n} cells, {pb.nnz} edges built in {time.perf_counter() - t0:.1f} s")
    tune = tune_packed(pb, use_cache=False, log=log)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    pb.run(steps)
    e1.record()
    torch.cuda.synchronize()
    pb.check()
    ms = e0.elapsed_time(e1)
    sim = pb.k * steps * lp.dt / 1000.0 / (ms / 1000.0)
    rate = float(pb.counts.sum().item()) / pb.n / (steps * lp.dt / 1000.0)
    log(f"packed: {steps} steps in one graph launch, {ms:.1f} ms ({ms * 1000 / steps:.1f} us/step): "
        f"{sim:.2f} instance-sim-s/wall-s, mean rate {rate:.1f} Hz")
    out = {"packed_subgraphs": pb.k, "packed_cells": pb.n, "packed_edges": pb.nnz, "packed_steps": steps,
           "packed_ms": round(ms, 2), "packed_us_per_step": round(ms * 1000 / steps, 2),
           "packed_instance_sim_s_per_wall_s": round(sim, 3), "packed_window": pb.window, "packed_grid": pb.grid,
           "packed_mean_rate_hz": round(rate, 2)}
    del pb
    torch.cuda.empty_cache()
    return out


def b1_run(conn, retina, lp, op, r, frames):
    """The plain single-fly path: FlyBrain(batch=1, cuda_graphs=True) with rollout r's inputs and movie, 10 frames to
    capture every delay-ring phase of a 100-step frame, then `frames` timed 10 ms frames."""
    import torch
    from flyhip.fly import FlyBrain
    log("B=1 plain path: setup")
    fb = FlyBrain(conn, retina, batch=1, device="cuda", seed=r.seed, lif_params=lp, optic_params=op, cuda_graphs=True)
    fb.set_poisson(r.input_idx, r.input_hz)
    fb.random_movie(r.movie_seed)
    for _ in range(10):
        fb.step(10.0)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(frames):
        fb.step(10.0)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    b1 = frames * 0.010 / (ms / 1000.0)
    log(f"B=1: {frames} frames in {ms:.1f} ms: {b1:.3f} sim-s/wall-s")
    del fb
    torch.cuda.empty_cache()
    return {"b1_sim_s_per_wall_s": round(b1, 4), "b1_ms_per_frame": round(ms / frames, 4)}


def copy_calibration(dev):
    """One device-to-device copy of 1 GiB: the kernel of known traffic the PMC byte counters are scaled by."""
    import torch
    x = torch.ones(2**28, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    log("copy calibration: 1 GiB read + 1 GiB written")


def attainable_copy_gbs(dev, reps=10):
    """The device-to-device copy bandwidth this GPU attains: 1 GiB copied `reps` times (after one untimed copy) by
    the library's streaming copy (launch_copy_stream) at 4, 8 and 16 workgroups per compute unit and by the HIP
    runtime's copy (torch copy_), timed with HIP events on the launch stream; the best of them, GB/s counting the
    bytes read and written."""
    import torch
    from flyhip import hip
    x = torch.ones(2**28, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    fn = hip.lib(dev).launch_copy_stream
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    runs = [lambda: y.copy_(x)]
    for per_cu in (4, 8, 16):
        def run(g=n_cu * per_cu):
            err = fn(x.data_ptr(), y.data_ptr(), x.numel() // 4, g, stream.cuda_stream)
            if err:
                raise RuntimeError(f"launch_copy_stream: HIP error {err}")
        runs.append(run)
    best = 0.0
    for run in runs:
        run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            run()
        e1.record(stream)
        e1.synchronize()
        best = max(best, 2 * x.numel() * 4 * reps / (e0.elapsed_time(e1) / 1e3) / 1e9)
    del x, y
    torch.cuda.empty_cache()
    return best


def abi_run(conn, retina, lp, op, batches=(1, 8), reps=20):
    """The reference's six-entry C ABI (flyverse/kernels/neural.cu:54, 80, 84, 105, 124, 141, called through flyhip.hip,
    the flyverse/cuda.py surface) on the full-scale graph (167,106 cells; the 89,390-unit optic lobe), at each batch
    size: every entry point timed with HIP events around one replay of a graph of `reps` calls (3 for launch_events),
    so the host's per-call work is not in the figure.
    launch_lif runs with Poisson forcing on 300 cells and an injected uniform per (fly, cell); launch_events scatters
    the spikes of 1.5 % of the cells through the Brain's CSC buffers and compact presynaptic list; launch_csr is the
    W_rr product; launch_means averages 100 groups of 50 cells.  LIF algorithmic bytes per call: per (fly, cell) ten
    4-byte fields read (v, g, drive, adapt, refrac, poisson_p, rnd, res, rate, spike_counts) and six written (v,
    refrac, spikes, rate, spike_counts, the spike slot), per cell active and std_u (8 bytes)."""
    import dataclasses
    import torch
    from flyhip import hip
    from flyhip.brain import Brain
    from flyhip.optic import OpticLobe, optic_matrices
    dev = "cuda"
    W_rr = hip.CSR(optic_matrices(conn, retina, op)["M_rr"], dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    out = {}

    def timed(fn, r):
        # r calls captured in one graph and replayed, so the figure is the GPU's time per call, not the host's
        fn()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(r):
                fn()
        graph.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(torch.cuda.current_stream())
        graph.replay()
        e1.record(torch.cuda.current_stream())
        e1.synchronize()
        del graph
        return e0.elapsed_time(e1) * 1000 / r

    for B in batches:
        brain = Brain(conn, dataclasses.replace(lp, event_driven=True), device=dev, batch=B)
        n = brain.n
        brain.set_poisson(torch.randperm(n, generator=torch.Generator().manual_seed(3))[:300].numpy(),
                          np.full(300, 150.0, np.float32))
        rnd = torch.rand(B, n, device=dev, generator=gen)
        x = (torch.rand(B, n, device=dev, generator=gen) < 0.015).float()
        optic = OpticLobe(conn, retina, op, device=dev, batch=B)
        m = optic.n_rate
        y, pr, sk = (torch.rand(B, m, device=dev, generator=gen) for _ in range(3))
        xr = torch.rand(B, m, device=dev, generator=gen)
        gptr = torch.arange(0, 101 * 50, 50, dtype=torch.int32, device=dev)
        gidx = torch.randperm(n, generator=torch.Generator().manual_seed(4))[:5000].to(torch.int32).to(dev)
        calls = {
            "lif": (lambda: hip.lif_update(brain, rnd), reps),
            "events": (lambda: hip.event_scatter(brain._out_ptr, brain._out_post, brain._out_w, x, brain.g,
                                                 brain._out_pre), 3),
            "csr_wrr": (lambda: W_rr.matvec_reference_shape(xr), reps),
            "optic_dr": (lambda: hip.optic_dr(optic.v, optic.b_vec, optic._cuda_dr), reps),
            "optic": (lambda: hip.optic_update(optic, y, pr, sk), reps),
            "means": (lambda: hip.group_means(gptr, gidx, brain.rate), reps),
        }
        for name, (fn, r) in calls.items():
            out[f"abi_{name}_b{B}_us"] = round(timed(fn, r), 2)
        lif_bytes = 4 * B * n * 16 + 8 * n
        out[f"abi_lif_b{B}_frac"] = round(lif_bytes / (out[f"abi_lif_b{B}_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
        del brain, optic
        torch.cuda.empty_cache()
    out["abi_cells"], out["abi_optic_units"], out["abi_wrr_nnz"] = int(conn.n), int(W_rr.shape[0]), int(W_rr.nnz)
    return out


def c0_run(cpu: bool, steps: int = 1000):
    """The plain-LIF workload in full (16,384 neurons, lognormal out-degree, 200 Poisson inputs at 150 Hz, dt 0.1 ms,
    1,000 steps; synth.plain_lif_graph, the configuration of the committed C0 golden): the plain HIP path (Brain,
    windows of up to 18 steps) timed with HIP events after 100 untimed steps, and with `cpu` the CPU path from rest
    on all host cores, the C oracle (OpenMP) and the reference's step in torch float32 ops, each for all `steps`."""
    import torch
    from flyhip import synth
    from flyhip.brain import Brain, LIFParams
    c, inputs, hz = synth.plain_lif_graph(0, n=16_384, n_inputs=200, hz=150.0)
    lp = LIFParams.shiu(dt=0.1)
    brain = Brain(c, lp, device="cuda", batch=1, seed=20240607)
    brain.set_poisson(inputs, hz)
    brain.step(100)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    brain.step(steps)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    out = {"c0_neurons": int(c.n), "c0_edges": int(c.W.nnz), "c0_steps": steps, "c0_gpu_ms": round(ms, 3),
           "c0_gpu_sim_s_per_wall_s": round(steps * lp.dt / ms, 4)}
    del brain
    if cpu:
        import scipy.sparse as sp
        sys.path.insert(0, str(ROOT))
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle
        from flyhip.brain import shaped_weights
        from helpers import oracle_fly
        from oracle.torch_ref import RefBrain
        cores = len(os.sched_getaffinity(0))
        oracle.set_threads(cores)
        om = oracle_fly(c, None, lp, None, 1, [20240607], None, with_optic=False)
        om.st.poisson_p[0, inputs] = (hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
        om.poisson_cells = np.asarray(inputs, np.int32)
        t0 = time.perf_counter()
        for _ in om.lif_steps(steps):
            pass
        wall_o = time.perf_counter() - t0
        torch.set_num_threads(cores)
        rb = RefBrain(sp.csr_matrix(shaped_weights(c, lp)), 1, lp, np.ones(c.n, np.float32))
        rb.poisson_p[0, torch.from_numpy(np.asarray(inputs, np.int64))] = torch.from_numpy(hz) * (lp.dt / 1000.0)
        gen = torch.Generator().manual_seed(0)
        t0 = time.perf_counter()
        for _ in range(steps):
            rb.step(torch.rand(1, c.n, generator=gen))
        wall_t = time.perf_counter() - t0
        out.update({"c0_cpu_cores": cores, "c0_cpu_oracle_openmp_ms": round(wall_o * 1e3, 1),
                    "c0_cpu_torch_reference_ops_ms": round(wall_t * 1e3, 1),
                    "c0_gpu_vs_cpu": round(min(wall_o, wall_t) * 1e3 / ms, 1)})
    return out


def cpu_baseline(conn, retina, lp, op, rollouts, oracle_frames, torch_frames):
    """The CPU path on every core of this host, B = len(rollouts) flies of the workload from rest:
      * the C oracle (oracle/, OpenMP over elements, rows and flies) for `oracle_frames` frames;
      * the reference's CPU step and optic frame in torch float32 ops (oracle.torch_ref: the event-driven synaptic
        input, the LIF of flyverse/brain.py:830-870, the optic frame of optic.py:563-604 with torch's own sparse
        products, torch.rand draws), torch.set_num_threads(cores), for `torch_frames` frames.
    Returns {kind: (aggregate sim-s/wall-s, wall s)} and the core count."""
    import torch
    import scipy.sparse as sp
    sys.path.insert(0, str(ROOT))
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle
    from flyhip.brain import shaped_weights
    from flyhip.optic import optic_matrices
    from helpers import oracle_fly
    from oracle.torch_ref import RefBrain, RefOptic
    cores = len(os.sched_getaffinity(0))
    B = len(rollouts)
    out = {}
    oracle.set_threads(cores)
    om = oracle_fly(conn, retina, lp, op, B, [r.seed for r in rollouts], [r.movie_seed for r in rollouts])
    for b, r in enumerate(rollouts):
        om.st.poisson_p[b, r.input_idx] = (r.input_hz * np.float32(lp.dt / 1000.0)).astype(np.float32)
    om.poisson_cells = np.unique(np.concatenate([r.input_idx for r in rollouts])).astype(np.int32)
    t0 = time.perf_counter()
    for _ in range(oracle_frames):
        om.frame(100)
    wall = time.perf_counter() - t0
    out["oracle_openmp"] = (oracle_frames * 0.010 * B / wall, wall)
    del om
    torch.set_num_threads(cores)
    m = optic_matrices(conn, retina, op)
    keep = np.ones(conn.n, np.float32)
    keep[m["rate_idx"]] = 0.0
    W = (sp.diags(keep) @ shaped_weights(conn, lp) @ sp.diags(keep)).tocsr()
    W.eliminate_zeros()
    rb = RefBrain(W, B, lp, keep)
    for b, r in enumerate(rollouts):
        rb.poisson_p[b, torch.from_numpy(np.asarray(r.input_idx, np.int64))] = torch.from_numpy(r.input_hz) * (lp.dt / 1000.0)
    ro = RefOptic(m, op, retina, B, conn.n, products="torch")
    gen = torch.Generator().manual_seed(0)
    t0 = time.perf_counter()
    for _ in range(torch_frames):
        rb.drive.copy_(ro.step_frame(torch.rand(B, retina.n_col, 4, generator=gen), rb.rate, 10.0))
        for _ in range(100):
            rb.step(torch.rand(B, conn.n, generator=gen))
    wall = time.perf_counter() - t0
    out["torch_reference_ops"] = (torch_frames * 0.010 * B / wall, wall)
    return out, cores


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60, help="timed 10 ms frames of the B=64 engine")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tune-warm", type=int, default=30, help="frames run before the tuning trials")
    ap.add_argument("--b1-frames", type=int, default=30)
    ap.add_argument("--cpu-frames", type=int, default=60, help="frames of the C oracle in the CPU baseline")
    ap.add_argument("--cpu-torch-frames", type=int, default=15, help="frames of the torch reference ops")
    ap.add_argument("--packed-steps", type=int, default=5000)
    ap.add_argument("--no-packed", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape of the B=64 engine")
    ap.add_argument("--b64-shape", default=None, metavar="FOLD,K,GRID_PER_CU",
                    help="pin the B=64 launch shape (fold mode, steps per launch, workgroups per CU); skips tuning")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-optic-training", action="store_true")
    ap.add_argument("--only", choices=("b64", "spmv", "b1", "optic-training", "packed", "c0", "abi"), default=None,
                    help="run one part alone (plus --copy-calibration) and print its figures: the short runs the "
                         "kernel traces and PMC counter passes profile")
    ap.add_argument("--only-b64", action="store_true", help="the same as --only b64")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device copy (the reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.only_b64:
        args.only = "b64"
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this workload shards by fly with no communication, and runs on more than "
              f"one GPU are not part of this build (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch
    from flyhip import synth
    from flyhip.batch import BatchBrain
    from flyhip.brain import LIFParams
    from flyhip.optic import OpticParams

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("building the full-scale graph ...")
    conn, retina, layout = synth.fly_graph(1, 1.0)
    lp = LIFParams.shiu(dt=0.1)
    op = OpticParams(tau_by_type=conn.rate_tau, baseline_by_type=conn.rate_baseline, pair_gain=[])
    rollouts = make_rollouts(layout, 64 + 4 * 64, seed=11)

    if args.only not in (None, "b64"):
        part = {"spmv": lambda: spmv_roofline(conn, retina, op), "b1": lambda: b1_run(conn, retina, lp, op, rollouts[0],
                                                                                        args.b1_frames),
                "optic-training": optic_training, "packed": lambda: packed_run(args.packed_steps, log),
                "c0": lambda: c0_run(not args.no_cpu), "abi": lambda: abi_run(conn, retina, lp, op)}[args.only]()
        if args.copy_calibration:
            copy_calibration(dev)
        print(json.dumps(part), flush=True)
        return

    # ---------------------------------------------------------------- B = 64, batched engine
    log("B=64 batched engine: setup")
    bb = BatchBrain(conn, retina, lp, op, device=dev)
    bb.submit(rollouts)
    bb.prepare(graph=True)
    tuned = None
    if args.b64_shape:
        fm, K_, gpc = (int(x) for x in args.b64_shape.split(","))
        tuned = {"fold_mode": fm, "steps_per_launch": K_, "grid_per_cu": gpc}
        bb.configure(**tuned)
    elif not args.no_tune:
        from flyhip.autotune import tune_batch
        tuned = tune_batch(bb, use_cache=False, log=log, warm=args.tune_warm)["config"]
        log(f"B=64 launch shape: {tuned}")
    bb.run(args.warmup)
    torch.cuda.synchronize()
    s0 = bb.simulated_steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    bb.run(args.steps)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    sim_steps = bb.simulated_steps - s0
    value = sim_steps * lp.dt / 1000.0 / (gpu_ms / 1000.0)
    log(f"B=64: {args.steps} frames, {sim_steps} fly-steps in {gpu_ms:.1f} ms (host wall {wall*1e3:.1f} ms): "
        f"{value:.3f} sim-s/wall-s; rollouts completed {bb.completed}")

    lif_ms, K = time_lif_launches(bb)
    alg_bytes, activity = lif_block_bytes(bb, K, 64)
    achieved = alg_bytes / (lif_ms / 1e3) / 1e9
    log(f"k_batch_lif: {lif_ms:.3f} ms per launch ({K:.3f} steps on average), activity {activity*100:.3f} % per "
        f"step, {alg_bytes/1e6:.1f} MB algorithmic -> {achieved:.0f} GB/s")
    sub_us, sub_bytes = time_optic_substeps(bb)
    sub_frac = sub_bytes / (sub_us * 1e-6) / 1e9 / HBM_PEAK_GBS
    log(f"k_optic_substep_fm: {sub_us:.1f} us, {sub_bytes / 1e6:.1f} MB algorithmic -> {sub_frac * 100:.1f} % of peak")
    if tuned is None:
        tuned = {"fold_mode": bb.fold_mode, "steps_per_launch": bb.steps_per_launch,
                 "grid_per_cu": bb.grid // torch.cuda.get_device_properties(dev).multi_processor_count}
    traffic, traffic_shape = measured_traffic(tuned)
    del bb
    torch.cuda.empty_cache()
    if args.only == "b64":
        if args.copy_calibration:
            copy_calibration(dev)
        print(json.dumps({"b64_sim_s_per_wall_s": round(value, 4), "lif_launch_ms": round(lif_ms, 4),
                          "optic_substep_b64_us": round(sub_us, 2), "optic_substep_b64_frac": round(sub_frac, 4)}),
              flush=True)
        return

    copy_gbs = attainable_copy_gbs(dev)
    log(f"attainable device copy: {copy_gbs:.0f} GB/s ({copy_gbs / HBM_PEAK_GBS * 100:.1f} % of the 8 TB/s peak)")
    spmv = spmv_roofline(conn, retina, op)
    spmv["spmv_wrr_b1_frac_attainable"] = round(spmv["spmv_wrr_b1_gbs"] / copy_gbs, 4)
    log(f"W_rr product, B=1: {spmv['spmv_wrr_b1_us']} us ({spmv['spmv_wrr_kernel']}), {spmv['spmv_wrr_b1_gbs']} GB/s "
        f"algorithmic ({spmv['spmv_wrr_b1_frac'] * 100:.1f} % of peak); all: "
        f"{ {k: v for k, v in spmv.items() if k.endswith('_us') and k != 'spmv_wrr_b1_us'} }")

    # ---------------------------------------------------------------- B = 1, plain kernels
    b1_part = b1_run(conn, retina, lp, op, rollouts[0], args.b1_frames)
    b1, b1_ms_frame = b1_part["b1_sim_s_per_wall_s"], b1_part["b1_ms_per_frame"]

    ol = None
    if not args.no_optic_training:
        log("optic-lobe training step (forward + adjoint), full scale ...")
        ol = optic_training()
        log(f"OL fwd+bwd: {ol['ol_fwd_bwd_ms_per_step']} ms per training step "
            f"({ol['ol_us_per_simulated_step']} us per simulated step)")
        torch.cuda.empty_cache()

    c0 = c0_run(not args.no_cpu)
    log(f"C0 plain LIF (16,384 neurons, 1,000 steps): GPU {c0['c0_gpu_ms']} ms"
        + (f", CPU oracle {c0['c0_cpu_oracle_openmp_ms']} ms, torch reference ops "
           f"{c0['c0_cpu_torch_reference_ops_ms']} ms on {c0['c0_cpu_cores']} cores" if "c0_cpu_cores" in c0 else ""))
    torch.cuda.empty_cache()

    packed = None
    if not args.no_packed:
        log("packed uneven subgraphs ...")
        packed = packed_run(args.packed_steps, log)

    if args.copy_calibration:
        copy_calibration(dev)

    # ---------------------------------------------------------------- CPU oracle on the host cores
    cpu = None
    if not args.no_cpu:
        log(f"CPU path, B=4 from rest: C oracle {args.cpu_frames} frames, torch reference ops {args.cpu_torch_frames} "
            f"frames ...")
        res, cores = cpu_baseline(conn, retina, lp, op, rollouts[:4], args.cpu_frames, args.cpu_torch_frames)
        best = max(res, key=lambda k: res[k][0])
        cpu = {"value": round(res[best][0], 6), "unit": "sim-s/wall-s", "cores": cores, "kind": "port",
               "sample": f"4 flies of the same workload from rest on all {cores} host cores ({cpu_model()}, "
                         f"os.cpu_count()={os.cpu_count()}): {best} (the faster); "
                         + "; ".join(f"{k} {v[0]:.5f} sim-s/wall-s ({v[1]:.1f} s wall, "
                                     f"{args.cpu_frames if k == 'oracle_openmp' else args.cpu_torch_frames} frames)"
                                     for k, v in res.items())
                         + "; the work grows linearly with the flies, so 64 flies run at this same aggregate rate",
               "by_kind": {k: round(v[0], 6) for k, v in res.items()}}
        log(f"CPU: " + ", ".join(f"{k} {v[0]:.5f} sim-s/wall-s ({v[1]:.1f} s)" for k, v in res.items()))

    line = {
        "metric": "batched sim-s/wall-s at 167k neurons (B=64) vs B=1; OL fwd+bwd ms/step",
        "value": round(value, 4), "unit": "sim-s/wall-s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(gpu_ms / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "full-scale fly (167,106 neurons, 25.4M edges, LIF dt 0.1 ms + 89,390-unit optic lobe), "
                               "B=64 batched, rollouts uniform 2,000-10,000 steps with lane refill; one step = one "
                               "10 ms frame",
                   "batch": 64, "neurons": int(conn.n), "edges": int(conn.W.nnz), "dt_ms": lp.dt,
                   "attainable_copy_gbs": round(copy_gbs, 1), "b1_sim_s_per_wall_s": b1, "ratio_b64_vs_b1": round(value / b1, 2),
                   "b1_ms_per_frame": b1_ms_frame, "spike_activity_per_step": activity,
                   "b64_launch_shape": tuned, "optic_substep_b64_us": round(sub_us, 2),
                   "optic_substep_b64_frac": round(sub_frac, 4), **spmv, **(ol or {}), **(packed or {}), **c0},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "attainable": round(copy_gbs, 1), "frac_attainable": round(achieved / copy_gbs, 4),
                     "kernel": "k_batch_lif", "launch_ms": round(lif_ms, 4), "steps_per_launch": round(K, 4),
                     "algorithmic_bytes_per_launch": int(alg_bytes),
                     "traffic_shape": traffic_shape,
                     "hbm_frac_measured": (round(traffic / (lif_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                                           if traffic else None),
                     "note": "achieved charges the per-step LIF state bytes once per step (K per launch) although "
                             "a launch keeps that state in registers across its steps; launch_ms brackets "
                             "k_batch_act and k_batch_lif; traffic is FETCH_SIZE*2 + WRITE_SIZE per launch from "
                             "profiles/hbm_traffic_k_batch_lif_b64.txt at traffic_shape, and hbm_frac_measured is "
                             "traffic / launch_ms / peak; attainable is this GPU's measured 1 GiB device copy "
                             "(read + write bytes)"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
