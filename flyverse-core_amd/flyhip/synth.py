"""Seeded synthetic connectomes and inputs for the brain-step workloads.

Every generator draws from numpy.random.default_rng(seed) only, so a (seed, sizes) pair names one graph.  Choices
that the workload statements leave open are fixed here and listed in each docstring.

Workloads:
  plain_lif_graph   16,384 LIF neurons; out-degree lognormal (mean 150, sigma 1), uniform targets, no self-loops;
                    counts geometric(0.2) clipped to [1, 200]; sign +1 for 70 % of presynaptic cells; 200 Poisson inputs.
  fly_graph         the full-scale fly: 167,106 neurons = 71,618 central LIF + 203 other sensory LIF + 5,895
                    photoreceptors + 89,390 optic-lobe rate units; 25.6 M edges in five blocks (LIF<-LIF 13.0 M,
                    rate<-rate 8.8 M, rate<-photoreceptor 0.3 M, rate<-LIF 1.0 M, LIF<-rate 2.5 M), in-degree
                    lognormal (sigma 1) within each block; counts floor(Pareto(1.2)) capped so the mean is ~4.85.
  fly_inputs        per fly: 300 Poisson inputs chosen among the 2,000 sensory (ORN/GRN-labelled) central cells at
                    U[0, 200] Hz.
  packed_subgraphs  lesion / region-extraction stand-ins: LIF graphs of log-uniform size with the fly graph's LIF
                    block statistics.
  packed_workload   those subgraphs with per-instance Philox seeds and 300 Poisson inputs at U[0, 200] Hz each.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import pandas as pd
import scipy.sparse as sp

from .connectome import Connectome, Retina

# full-scale fly composition (the MaleCNS counts of the reference, README.md:58-60)
N_CENTRAL, N_OTHER_SENSORY, N_PR, N_RATE = 71_618, 203, 5_895, 89_390
N_FLY = N_CENTRAL + N_OTHER_SENSORY + N_PR + N_RATE          # 167,106
N_INPUT_POOL = 2_000                                           # ORN/GRN-labelled central cells
BLOCK_EDGES = {"lif": 13_000_000, "rr": 8_800_000, "rp": 300_000, "rs": 1_000_000, "sr": 2_500_000}


def pareto_cap(alpha: float = 1.2, mean: float = 4.85) -> int:
    """Cap K such that E[min(floor(X), K)] = sum_{k<=K} k^-alpha reaches `mean` for X ~ Pareto(alpha, x_m = 1)."""
    s, k = 0.0, 0
    while s < mean:
        k += 1
        s += k ** -alpha
    return k


PARETO_CAP = pareto_cap()


def power_law_counts(rng, size, alpha=1.2, cap=PARETO_CAP) -> np.ndarray:
    u = rng.random(size)
    x = np.floor((1.0 - u) ** (-1.0 / alpha))
    return np.minimum(x, cap).astype(np.float32)


def _block(rng, rows: np.ndarray, cols: np.ndarray, total: int, sigma: float = 1.0):
    """Edges (row, col) with lognormal in-degree (mean total/len(rows)) and uniform distinct presynaptic columns."""
    m = total / max(len(rows), 1)
    deg = np.rint(rng.lognormal(math.log(m) - sigma * sigma / 2, sigma, len(rows))).astype(np.int64)
    deg = np.clip(deg, 0, len(cols))
    r = np.repeat(rows.astype(np.int64), deg)
    c = cols[rng.integers(0, len(cols), size=len(r))].astype(np.int64)
    return r, c


def _assemble(n: int, r: np.ndarray, c: np.ndarray, counts_fn, sign: np.ndarray) -> sp.csr_matrix:
    key = np.unique(r * n + c)                     # sorted: post-major, presynaptic ascending; duplicates dropped
    post = (key // n).astype(np.int32)
    pre = (key % n).astype(np.int32)
    counts = counts_fn(len(key))
    data = (counts * sign[pre]).astype(np.float32)
    indptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(post, minlength=n), out=indptr[1:])
    return sp.csr_matrix((data, pre, indptr), shape=(n, n))


def _neurons(types, superclass) -> pd.DataFrame:
    n = len(types)
    return pd.DataFrame({"bodyId": np.arange(n, dtype=np.int64) + 10_000, "type": types, "superclass": superclass})


def plain_lif_graph(seed: int = 0, n: int = 16_384, n_inputs: int = 200, hz: float = 150.0):
    """The plain-LIF workload: returns (Connectome, input indices, input Hz).  Duplicated (post, pre) draws are
    summed as Brain.set_weights does (flyverse/brain.py:530); self-loops are excluded by construction."""
    rng = np.random.default_rng(seed)
    sign = np.where(rng.random(n) < 0.7, 1.0, -1.0).astype(np.float32)
    deg = np.rint(rng.lognormal(math.log(150.0) - 0.5, 1.0, n)).astype(np.int64)
    deg = np.clip(deg, 1, n - 1)
    pre = np.repeat(np.arange(n, dtype=np.int64), deg)
    post = rng.integers(0, n - 1, size=len(pre))
    post = post + (post >= pre)                     # uniform over the other n-1 cells
    counts = np.clip(rng.geometric(0.2, size=len(pre)), 1, 200).astype(np.float32)
    W = sp.coo_matrix((counts * sign[pre], (post, pre)), shape=(n, n)).tocsr()   # sums duplicates
    W.sort_indices()
    types = np.array([f"LIF{i}" for i in range(n)], dtype=object)
    c = Connectome(_neurons(types, np.full(n, "cb_intrinsic", dtype=object)), W)
    inputs = np.sort(rng.choice(n, n_inputs, replace=False))
    return c, inputs, np.full(n_inputs, hz, np.float32)


@dataclass
class FlyLayout:
    n: int
    central: np.ndarray
    sensory: np.ndarray
    pr: np.ndarray
    rate: np.ndarray
    input_pool: np.ndarray


def fly_layout(scale: float = 1.0) -> FlyLayout:
    nc, ns, npr, nr = (max(1, int(round(x * scale))) for x in (N_CENTRAL, N_OTHER_SENSORY, N_PR, N_RATE))
    n = nc + ns + npr + nr
    central = np.arange(nc)
    sensory = np.arange(nc, nc + ns)
    pr = np.arange(nc + ns, nc + ns + npr)
    rate = np.arange(nc + ns + npr, n)
    pool = central[:max(1, int(round(N_INPUT_POOL * scale)))]
    return FlyLayout(n, central, sensory, pr, rate, pool)


def fly_graph(seed: int = 1, scale: float = 1.0):
    """The full-scale fly (scale 1) or a proportionally shrunk one: returns (Connectome, Retina, FlyLayout).

    Rate units get one cell type each ("OL<i>") so per-unit time constants and operating points are expressed
    through the reference's per-type maps (OpticParams.tau_by_type / baseline_by_type, flyverse/optic.py:343-347):
    tau ~ lognormal around 20 ms (sigma 0.5), operating point b ~ N(0, 0.1)."""
    rng = np.random.default_rng(seed)
    L = fly_layout(scale)
    n = L.n
    spk = np.concatenate([L.central, L.sensory])
    sign = np.where(rng.random(n) < 0.7, 1.0, -1.0).astype(np.float32)
    parts = [
        _block(rng, spk, spk, int(BLOCK_EDGES["lif"] * scale)),
        _block(rng, L.rate, L.rate, int(BLOCK_EDGES["rr"] * scale)),
        _block(rng, L.rate, L.pr, int(BLOCK_EDGES["rp"] * scale)),
        _block(rng, L.rate, spk, int(BLOCK_EDGES["rs"] * scale)),
        _block(rng, spk, L.rate, int(BLOCK_EDGES["sr"] * scale)),
    ]
    r = np.concatenate([p[0] for p in parts])
    c = np.concatenate([p[1] for p in parts])
    del parts
    W = _assemble(n, r, c, lambda k: power_law_counts(rng, k), sign)
    del r, c
    types = np.empty(n, dtype=object)
    superclass = np.empty(n, dtype=object)
    types[L.central] = "CB"
    types[L.input_pool] = [("ORN" if i % 2 == 0 else "GRN") for i in range(len(L.input_pool))]
    superclass[L.central] = "cb_intrinsic"
    superclass[L.input_pool] = "cb_sensory"
    types[L.sensory] = "OLS"
    superclass[L.sensory] = "ol_sensory"
    types[L.pr] = "R1-R6"
    superclass[L.pr] = "ol_sensory"
    types[L.rate] = [f"OL{i}" for i in range(len(L.rate))]
    superclass[L.rate] = "ol_intrinsic"
    conn = Connectome(_neurons(types, superclass), W)
    sens = rng.random((len(L.pr), 4)).astype(np.float32)
    sens /= sens.sum(axis=1, keepdims=True)
    retina = Retina(pr_index=L.pr, pr_sens=sens.astype(np.float32))
    tau = np.exp(rng.normal(math.log(20.0), 0.5, len(L.rate)))
    base = rng.normal(0.0, 0.1, len(L.rate)).astype(np.float32)
    conn.rate_tau = {f"OL{i}": float(tau[i]) for i in range(len(L.rate))}
    conn.rate_baseline = {f"OL{i}": float(base[i]) for i in range(len(L.rate))}
    return conn, retina, L


def fly_inputs(seed: int, layout: FlyLayout, n_inputs: int = 300, hz_max: float = 200.0):
    """One fly's sensory drive: (indices, Hz) of its Poisson-forced cells."""
    rng = np.random.default_rng(seed)
    k = min(n_inputs, len(layout.input_pool))
    idx = np.sort(rng.choice(layout.input_pool, k, replace=False))
    return idx, rng.uniform(0.0, hz_max, k).astype(np.float32)


def packed_subgraphs(seed: int = 3, count: int = 128, n_min: int = 5_000, n_max: int = 80_000,
                     mean_in_degree: float = 167.0, which=None):
    """`count` LIF subgraphs with sizes log-uniform in [n_min, n_max] and the fly LIF block's statistics (in-degree
    lognormal sigma 1 around `mean_in_degree`, power-law counts, 70 % excitatory presynaptic cells).  Sizes come from
    default_rng(seed), instance k's edges from default_rng([seed, k]), so any instance can be built alone.  Returns
    the list of per-instance (post, pre) signed-count CSR matrices, of instances `which` (default: all)."""
    sizes = np.exp(np.random.default_rng(seed).uniform(math.log(n_min), math.log(n_max), count)).astype(np.int64)
    mats = []
    for k in (range(count) if which is None else which):
        rng = np.random.default_rng([seed, k])
        nk = int(sizes[k])
        cells = np.arange(nk)
        sign = np.where(rng.random(nk) < 0.7, 1.0, -1.0).astype(np.float32)
        r, c = _block(rng, cells, cells, int(mean_in_degree * nk))
        mats.append(_assemble(nk, r, c, lambda m: power_law_counts(rng, m), sign))
    return mats


def packed_workload(seed: int = 3, count: int = 128, n_min: int = 5_000, n_max: int = 80_000,
                    mean_in_degree: float = 153.0, input_seed: int = 12, n_inputs: int = 300, max_hz: float = 200.0,
                    which=None):
    """The uneven-size batch: `count` subgraphs of log-uniform size (packed_subgraphs), each with its own Philox seed
    and `n_inputs` distinct Poisson input cells at U[0, max_hz] Hz (drawn from default_rng([input_seed, k]) for
    instance k).  Returns (mats, inputs, seeds) for instances `which` (default: all): inputs[j] = (sorted cell
    indices, float32 rates), seeds a uint64 array."""
    ks = list(range(count) if which is None else which)
    mats = packed_subgraphs(seed, count, n_min, n_max, mean_in_degree=mean_in_degree, which=ks)
    inputs, seeds = [], []
    for k, m in zip(ks, mats):
        rng = np.random.default_rng([input_seed, k])
        inputs.append((np.sort(rng.choice(m.shape[0], n_inputs, replace=False)),
                       rng.uniform(0, max_hz, n_inputs).astype(np.float32)))
        seeds.append(rng.integers(1, 2**62, dtype=np.uint64))
    return mats, inputs, np.array(seeds, np.uint64)
