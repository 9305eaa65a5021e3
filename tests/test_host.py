"""Host-side logic on the CPU: the C ABI library exports exactly what include/flyhip.h declares, the graph builders
and weight shaping produce what the reference's formulas give, the packed CSR is the block-diagonal of its
instances, and every GPU entry refuses a device it cannot serve instead of falling back."""
import ctypes as ct
import math
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp

from flyhip import hip, synth
from flyhip.brain import Brain, LIFParams, shaped_weights
from flyhip.connectome import Connectome
from flyhip.fly import StepResult, budget_chunk
from flyhip.packed import PackedBrains, block_schedule, pack_instances

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    text = (ROOT / "include" / "flyhip.h").read_text()
    return sorted(set(re.findall(r"^int\s+(\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_header_symbol():
    if not hip.LIB_PATH.exists():
        from flyhip import build
        build.build()
    L = ct.CDLL(str(hip.LIB_PATH))
    names = header_symbols()
    assert len(names) >= 38
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"declared in include/flyhip.h but not exported: {missing}"
    # and nothing exported that the header does not declare
    out = subprocess.run(["nm", "-D", "--defined-only", str(hip.LIB_PATH)], capture_output=True, text=True).stdout
    exported = sorted({f.split()[-1] for f in out.splitlines() if re.search(r" T (launch_|packed_)", f)})
    assert exported == names


def test_reference_signatures_are_registered():
    """The six reference entry points keep flyverse/cuda.py:73-75's argument lists."""
    ptrs, i32, f32, stream = ct.POINTER(ct.c_void_p), ct.c_int, ct.c_float, ct.c_void_p
    expected = {"launch_lif": [ptrs, i32, i32, i32, stream], "launch_optic": [ptrs, i32, i32, f32, f32, f32, stream],
                "launch_csr": [ptrs, i32, i32, i32, i32, stream], "launch_events": [ptrs, i32, i32, i32, i32, stream],
                "launch_means": [ptrs, i32, i32, i32, i32, stream], "launch_optic_dr": [ptrs, i32, i32, stream]}
    assert {name: hip.SIGNATURES[name] for name in expected} == expected


def small_connectome(rng, n=40, density=0.2):
    W = sp.random(n, n, density=density, random_state=rng, format="csr", dtype=np.float32)
    W.data = np.ceil(W.data * 9).astype(np.float32) * np.where(rng.random(W.nnz) < 0.7, 1, -1).astype(np.float32)
    types = np.array([f"T{i % 5}" for i in range(n)], dtype=object)
    sc = np.array(["cb_intrinsic" if i % 3 else "cb_sensory" for i in range(n)], dtype=object)
    return Connectome(pd.DataFrame({"bodyId": np.arange(n) + 100, "type": types, "superclass": sc}), W)


def test_shaped_weights_follow_the_reference_formulas():
    """cap, same-type damping, fan-in scale and w_syn in flyverse/brain.py:302-338, 407-426 order, checked in
    float64 against a direct restatement per entry."""
    rng = np.random.default_rng(4)
    c = small_connectome(rng)
    p = LIFParams(conn_cap=5.0, same_type_gain=0.5, input_norm_alpha=0.5, input_norm_ref=20.0, path_gain=[],
                  type_path_gain=[], w_syn=0.275)
    W = shaped_weights(c, p).toarray().astype(np.float64)
    A = c.W.toarray().astype(np.float64)
    A = np.sign(A) * np.minimum(np.abs(A), 5.0)
    t = c.neurons.type.to_numpy()
    same = (t[:, None] == t[None, :])
    A = np.where(same, A * 0.5, A)
    tot = np.abs(A).sum(axis=1)
    scale = np.clip((20.0 / np.maximum(tot, 1.0)) ** 0.5, 0.02, 1.0)
    A = A * scale[:, None] * 0.275
    assert np.allclose(W, A, rtol=1e-6, atol=1e-7)
    # the Shiu setting is counts x w_syn exactly
    Ws = shaped_weights(c, LIFParams.shiu()).toarray()
    assert np.array_equal(Ws, (c.W.toarray() * np.float32(LIFParams.shiu().w_syn)).astype(np.float32))


def test_plain_lif_graph_statistics():
    """The plain-LIF workload: lognormal out-degree around 150, geometric(0.2) counts in [1, 200], signs fixed per
    presynaptic cell with 70 % excitatory, no self-loops."""
    c, inputs, hz = synth.plain_lif_graph(0)
    W = c.W.tocsc()
    n = c.n
    assert n == 16_384 and len(inputs) == 200 and np.all(hz == 150.0)
    assert W.diagonal().sum() == 0
    outdeg = np.diff(W.indptr)
    assert 120 < outdeg.mean() < 160                       # duplicates merged: slightly below the drawn 150
    signs = np.array([np.sign(W.data[W.indptr[j]:W.indptr[j + 1]]).mean() if outdeg[j] else 1 for j in range(n)])
    assert set(np.unique(signs)) <= {-1.0, 1.0}            # one sign per presynaptic cell
    assert abs((signs > 0).mean() - 0.7) < 0.02
    counts = np.abs(W.data)
    assert counts.min() >= 1 and abs(counts.mean() - 5.0) < 0.3   # geometric(0.2) has mean 5 before merging


def test_fly_graph_layout_at_small_scale():
    conn, retina, L = synth.fly_graph(1, 0.02)
    assert conn.n == L.n == len(L.central) + len(L.sensory) + len(L.pr) + len(L.rate)
    assert len(retina.pr_index) == len(L.pr) and np.allclose(retina.pr_sens.sum(axis=1), 1.0, atol=1e-6)
    W = conn.W.tocsr()
    W.sort_indices()
    assert np.all(np.diff(W.indptr) >= 0) and W.has_sorted_indices
    assert set(conn.rate_tau) == {f"OL{i}" for i in range(len(L.rate))}


def test_pack_instances_is_the_block_diagonal():
    mats = synth.packed_subgraphs(2, 5, 50, 400, mean_in_degree=12.0)
    p = LIFParams.shiu()
    indptr, indices, data = pack_instances(mats, p)
    blocks = []
    for m in mats:
        n = m.shape[0]
        c = Connectome(pd.DataFrame({"bodyId": np.arange(n), "type": [""] * n, "superclass": ["cb_intrinsic"] * n}),
                       sp.csr_matrix(m))
        blocks.append(shaped_weights(c, p))
    ref = sp.block_diag(blocks, format="csr", dtype=np.float32)
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(indptr, ref.indptr) and np.array_equal(indices, ref.indices)
    assert np.array_equal(data, ref.data)


def test_block_schedule_visits_every_block_once_and_deals_instances_to_xcds():
    rng = np.random.default_rng(4)
    sizes = rng.integers(40, 900, 24)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    n = int(offsets[-1])
    deg = rng.integers(0, 700, n)
    indptr = np.concatenate([[0], np.cumsum(deg)])
    order = block_schedule(indptr, offsets)
    words = (n + 63) // 64
    assert order.dtype == np.int32 and sorted(order.tolist()) == list(range(words))
    # while every group has blocks left, position i takes a block of group (i // 4) % 8, and a group's blocks come
    # from few instances (each instance lives in one group)
    binst = np.searchsorted(offsets, order.astype(np.int64) * 64, side="right") - 1
    owner = {}
    for i, k in enumerate(binst[: 8 * 4 * 2]):
        owner.setdefault(int(k), set()).add((i // 4) % 8)
    assert all(len(g) == 1 for g in owner.values())


def test_gpu_entry_points_refuse_the_cpu():
    rng = np.random.default_rng(0)
    c = small_connectome(rng)
    with pytest.raises(ValueError):
        Brain(c, LIFParams.shiu(), device="cpu")
    with pytest.raises(ValueError):
        PackedBrains([c.W], device="cpu")
    with pytest.raises(ValueError):
        hip.use("cpu", explicit=True)
    assert hip.use("cpu") is False


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


# cap, same-type damping and a fan-in normalisation that actually scales these small graphs
SHAPING = LIFParams(conn_cap=5.0, same_type_gain=0.5, input_norm_alpha=0.5, input_norm_ref=20.0, path_gain=[],
                    type_path_gain=[], w_syn=0.275)


def test_subset_is_the_induced_subgraph_with_the_original_normalisation():
    rng = np.random.default_rng(9)
    c = small_connectome(rng, n=90)
    p = SHAPING
    idx = rng.permutation(c.n)[: c.n // 3]
    sub = c.subset(idx)
    assert sub.n == len(idx) and sub.reference is c
    assert (sub.W != c.W[idx][:, idx]).nnz == 0
    assert np.array_equal(sub.neurons.bodyId.to_numpy(), c.neurons.bodyId.to_numpy()[idx])
    assert np.array_equal(sub.index_of(sub.neurons.bodyId.to_numpy()[::-1]), np.arange(sub.n)[::-1])
    # every rule is per edge and the fan-in scale of a row comes from the original graph, so the shaped subset is
    # the induced subgraph of the shaped original
    full = shaped_weights(c, p)
    assert abs(shaped_weights(sub, p) - full[idx][:, idx]).max() == 0
    # a nested subset keeps the same original
    nested = sub.subset(np.arange(0, sub.n, 2))
    assert nested.reference is c
    assert abs(shaped_weights(nested, p) - full[idx[::2]][:, idx[::2]]).max() == 0
    # standing alone the subset would normalise by its own (smaller) fan-in
    alone = Connectome(sub.neurons.copy(), sub.W)
    assert abs(shaped_weights(alone, p) - full[idx][:, idx]).max() > 0
    with pytest.raises(ValueError):
        c.subset([0, 0])
    with pytest.raises(ValueError):
        c.subset(np.ones(c.n + 1, bool))
    assert sub.subset(np.ones(sub.n, bool)).n == sub.n


def test_cache_round_trip(tmp_path):
    rng = np.random.default_rng(10)
    c = small_connectome(rng, n=60)
    sub = c.subset(np.arange(5, c.n - 5))
    sub.save(tmp_path / "cache")
    back = Connectome.load_cache(tmp_path / "cache")
    assert (back.W != sub.W).nnz == 0 and back.W.dtype == np.float32
    assert back.neurons.equals(sub.neurons)
    assert (back.reference.W != c.W).nnz == 0
    assert abs(shaped_weights(back, SHAPING) - shaped_weights(sub, SHAPING)).max() == 0
    c.save(tmp_path / "full")
    assert not (tmp_path / "full" / "reference_W.npz").exists()
    assert Connectome.load_cache(tmp_path / "full").reference is not c


def test_packed_instances_accept_subsets():
    rng = np.random.default_rng(11)
    c = small_connectome(rng, n=90)
    p = SHAPING
    parts = [c.subset(np.arange(0, 40)), c.subset(np.arange(40, c.n))]
    indptr, indices, data = pack_instances(parts, p)
    ref = sp.block_diag([shaped_weights(x, p) for x in parts], format="csr", dtype=np.float32)
    ref.sum_duplicates()
    ref.sort_indices()
    assert np.array_equal(indptr, ref.indptr) and np.array_equal(indices, ref.indices)
    assert np.array_equal(data, ref.data)


def test_budget_chunk_follows_measured_costs():
    """The wall-budget chunk choice of flyverse/fly.py:640-648: one step before any measurement, then the largest
    candidate whose measured (or per-step extrapolated) cost fits 1.1x within the remaining time."""
    cands = [1, 2, 4, 8, 16, 32, 64, 100]
    assert budget_chunk({}, cands, 0.5) == 1
    assert budget_chunk({1: 2.0}, cands, 50.0) == 16              # 2 ms/step: 16 * 2.2 = 35.2 <= 50 < 32 * 2.2
    assert budget_chunk({1: 2.0, 32: 20.0}, cands, 50.0) == 32    # a measured chunk beats the extrapolation
    assert budget_chunk({4: 4.0}, cands, 4.4) == 4                # smallest measured chunk sets the per-step cost
    assert budget_chunk({4: 4.0}, cands, 1.0) == 0                # nothing fits
    assert budget_chunk({1: 0.01}, cands, 1e9) == 100             # the whole frame is the largest candidate
    r = StepResult(50.0, 25.0, 500)
    assert r.time_dilation == 2.0 and StepResult(1.0, 0.0, 10).time_dilation == 0.0


@pytest.mark.parametrize("norm,out_norm", [("l1", "l1"), ("l2", "l2"), ("l2", "l1")])
def test_optic_partition_normalisation_and_pair_gains(norm, out_norm):
    """optic_matrices against a dense float64 restatement of flyverse/optic.py:229-273: the three cell classes, each
    row divided by max(its L1 or L2 input total, 1), the pair gains applied in rule order to the rate->rate and
    rate->spiking blocks (anchored regular expressions, double factors), then float32."""
    from flyhip.optic import OpticParams, optic_matrices
    conn, retina, layout = synth.fly_graph(2, 0.01)
    rules = [("OL5", "OL", 1.5), ("OL", "OL59", 0.7), ("CB", "OL", 1.25), ("OL", "CB", 0.5), ("OL1", "OL1", 3.0)]
    op = OpticParams(tau_by_type=conn.rate_tau, baseline_by_type=conn.rate_baseline, pair_gain=rules, norm=norm,
                     out_norm=out_norm)
    m = optic_matrices(conn, retina, op)
    nrn = conn.neurons
    types = nrn.type.fillna("").to_numpy()
    is_pr_type = nrn.type.isin(["R1-R6", "R7p", "R7y", "R7d", "R7_unclear", "R8p", "R8y", "R8d", "R8_unclear",
                                "R7R8_unclear"]).to_numpy()
    rate = [i for i in range(conn.n) if nrn.superclass[i] == "ol_intrinsic" and not is_pr_type[i]]
    placed = set(retina.pr_index.tolist())
    spk = [i for i in range(conn.n) if not is_pr_type[i] and i not in rate and i not in placed]
    assert m["rate_idx"].tolist() == rate and m["spk_idx"].tolist() == spk
    D = conn.W.toarray().astype(np.float64)
    l1 = D / np.maximum(nrn.in_syn.to_numpy(), 1.0)[:, None]
    l2 = D / np.maximum(nrn.in_syn_l2.to_numpy(), 1.0)[:, None]
    into_rate = l2 if norm == "l2" else l1

    def gained(M, post_t, pre_t):
        M = M.copy()
        for pre_re, post_re, f in rules:
            rows = np.array([re.match(post_re, t) is not None for t in post_t])
            cols = np.array([re.match(pre_re, t) is not None for t in pre_t])
            sel = np.outer(rows, cols) & (M != 0)
            M[sel] = M[sel] * f
        return M.astype(np.float32)

    rt, st = types[rate], types[spk]
    want = {"M_rr": gained(into_rate[np.ix_(rate, rate)], rt, rt),
            "M_rs": into_rate[np.ix_(rate, spk)].astype(np.float32),
            "M_sr": gained((l2 if out_norm == "l2" else l1)[np.ix_(spk, rate)], st, rt),
            "M_rp": into_rate[np.ix_(rate, retina.pr_index)].astype(np.float32)}
    assert not np.array_equal(want["M_rr"], into_rate[np.ix_(rate, rate)].astype(np.float32))   # the rules apply
    assert not np.array_equal(want["M_sr"], (l2 if out_norm == "l2" else l1)[np.ix_(spk, rate)].astype(np.float32))
    for name, W in want.items():
        got = m[name]
        assert np.array_equal(got.toarray(), W), name
        assert got.has_canonical_format and got.dtype == np.float32
    assert np.array_equal(m["a"], np.exp(-op.dt_ms / np.array([conn.rate_tau.get(t, op.tau_ms) for t in rt],
                                                               np.float32)).astype(np.float32))


def test_frame_clock_takes_whole_steps_and_carries_the_rest():
    """floor((carry + frame + 1e-9) / dt) steps and a nonnegative carry (flyverse/fly.py:421-425): 1 ms of 0.1 ms
    steps is 10 steps (floor division of floats would give 9), 0.25 ms frames of 0.1 ms steps alternate 2, 3."""
    from flyhip.clock import split_frame, whole_steps
    assert whole_steps(1.0, 0.1) == 10 and 1.0 // 0.1 == 9.0
    carry, taken = 0.0, []
    for _ in range(4):
        n, carry = split_frame(carry, 0.25, 0.1)
        taken.append(n)
        assert carry >= 0.0
    assert taken == [2, 3, 2, 3]
    assert split_frame(0.0, 0.0, 0.5) == (0, 0.0)
    n, carry = split_frame(0.0, 0.3, 0.1)            # 0.1 + 0.1 + 0.1 rounds just under 0.3: the allowance keeps 3
    assert n == 3 and carry == 0.0


def test_select_criteria_and_dictionary_selections():
    """Connectome.select (flyverse/connectome.py:604-616): scalars by equality, collections by membership, '~'
    strings as regular expressions searched anywhere with missing values as empty; all criteria must hold.  A
    dictionary passed to indices() or subset() selects the same rows."""
    neurons = pd.DataFrame({"bodyId": [10, 11, 12, 13, 14, 15],
                            "type": ["DNa01", "DNa02", "DNa03", None, "xDNa01", "LC4"],
                            "somaSide": ["L", "R", "L", "L", "R", "L"],
                            "superclass": ["descending", "descending", "descending", "ol_intrinsic", "descending",
                                           "visual_projection"]})
    c = Connectome(neurons, sp.csr_matrix((6, 6), dtype=np.float32))
    assert c.select(somaSide="L").tolist() == [0, 2, 3, 5]
    assert c.select(type="~^DNa0[12]").tolist() == [0, 1]          # anchored at the start of the text
    assert c.select(type="~DNa01").tolist() == [0, 4]              # found anywhere, as str.contains does
    assert c.select(type="~^$").tolist() == [3]                    # the missing type reads as ""
    assert c.select(type=["LC4", "DNa03"]).tolist() == [2, 5]
    assert c.select(type={"LC4"}, somaSide="L").tolist() == [5]
    assert c.select(type="~^DNa0[12]", somaSide="L").tolist() == [0]
    assert c.select(superclass="nothing").tolist() == []
    assert c.select().tolist() == list(range(6))
    assert c.indices({"somaSide": "R"}).tolist() == [1, 4]
    sub = c.subset({"superclass": "descending", "somaSide": "L"})
    assert sub.neurons.bodyId.tolist() == [10, 12]
    with pytest.raises(KeyError):
        c.select(no_such_column=1)


def test_batch_launch_shapes():
    """The launch shapes the batched engine accepts: modes 0-6, 1..n_delay steps, mode 6 only on the slot-major
    ring with at most 8 steps (its activity bits are one byte per row)."""
    from flyhip.batch import check_launch_shape
    for mode in range(7):
        check_launch_shape(mode, 6, 18, "slot")
    check_launch_shape(5, 18, 18, "row")
    check_launch_shape(6, 8, 18, "slot")
    for bad in [(7, 6, 18, "slot"), (-1, 6, 18, "slot"), (5, 0, 18, "slot"), (5, 19, 18, "slot"),
                (6, 9, 18, "slot"), (6, 6, 18, "row")]:
        with pytest.raises(ValueError):
            check_launch_shape(*bad)


def test_tuner_cache_key_ring_layout_and_refused_shapes(monkeypatch):
    """The batched tuner's cache key names the spike ring's layout; a cached shape the engine would refuse (fold mode
    6 on a row-major ring) is not used; a search space with no acceptable shape is an error, raised before any GPU
    work."""
    from types import SimpleNamespace
    from flyhip import autotune
    monkeypatch.setattr(autotune, "_device_tag", lambda device: "testgpu:256")
    fake = lambda ring: SimpleNamespace(device="cuda", n_rows=10, W_lif=SimpleNamespace(nnz=5), n_rate=3, n_delay=18,
                                        ring_layout=ring)
    slot, row = fake("slot"), fake("row")
    assert autotune.batch_key(slot) != autotune.batch_key(row) and autotune.batch_key(row).endswith("ring-row")
    table = {autotune.batch_key(row): {"config": {"fold_mode": 6, "steps_per_launch": 6, "grid_per_cu": 128}}}
    assert autotune.cached_config(table, autotune.batch_key(row), 18, "row") is None
    table[autotune.batch_key(slot)] = {"config": {"fold_mode": 6, "steps_per_launch": 6, "grid_per_cu": 128}}
    assert autotune.cached_config(table, autotune.batch_key(slot), 18, "slot") is not None
    assert autotune.cached_config({"k": {"config": {}}}, "k", 18, "slot") is None
    shapes = autotune.candidate_shapes((5, 6), (4, 9), (64,), 18, "slot")
    assert {(s_["fold_mode"], s_["steps_per_launch"]) for s_ in shapes} == {(5, 4), (5, 9), (6, 4)}
    small = SimpleNamespace(**vars(fake("slot")) | {"n_delay": 3})
    with pytest.raises(ValueError, match="no launch shape to try"):
        autotune.tune_batch(small, steps=(4, 5, 6), use_cache=False)


def test_panel_blocks_cover_the_rows_within_the_register_budget():
    """The row blocks of the panel product: consecutive, covering every row once, at most 3 * 1024 chunks each, at
    least one block per compute unit when the rows allow, and no blocks when one row alone is too long."""
    rng = np.random.default_rng(3)
    cap = hip.PANEL_THREADS * hip.PANEL_CHUNKS
    for counts, n_cu in ((rng.integers(0, 40, 90000), 256), (rng.integers(1, 300, 5000), 256),
                         (np.array([3000, 3072, 1, 0, 2000]), 2), (np.zeros(7, np.int64), 256), (np.ones(3), 256)):
        starts = hip.panel_blocks(counts, n_cu)
        assert starts[0] == 0 and starts[-1] == len(counts) and np.all(np.diff(starts) > 0)
        sums = np.add.reduceat(counts, starts[:-1]) if len(counts) else np.zeros(0)
        assert sums.max(initial=0) <= cap
        assert len(starts) - 1 >= min(n_cu, len(counts))
    assert hip.panel_blocks(np.array([5, cap + 1, 5]), 256) is None
