"""Launch-shape autotuner for the batched engine and the packed-subgraph stepper.

Every shape it chooses between performs the same additions in the same order (the fold modes, the steps per
launch, the grid size, the rounds per pull), so results are bit-identical whichever wins; only speed differs.  The
winner is cached in a JSON file keyed by the graph's shape, so a rerun on the same workload skips the search.
"""
from __future__ import annotations

import itertools
import json
import os
from pathlib import Path

import torch

CACHE = Path(os.environ.get("FLYHIP_TUNE_CACHE", Path.home() / ".cache" / "flyhip" / "tune.json"))


def _load() -> dict:
    try:
        return json.loads(CACHE.read_text())
    except (OSError, ValueError):
        return {}


def _save(table: dict) -> None:
    try:
        CACHE.parent.mkdir(parents=True, exist_ok=True)
        tmp = CACHE.with_suffix(f".tmp{os.getpid()}")
        tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
        tmp.replace(CACHE)
    except OSError:
        pass


def _device_tag(device) -> str:
    dev = torch.cuda.get_device_properties(device)
    return f"{dev.name}:{dev.multi_processor_count}"


def batch_key(bb) -> str:
    """Cache key of a batched engine's tuning: the device, the graph's shape and the spike ring's layout (fold mode 6
    exists only on the slot-major ring)."""
    return (f"batch:{_device_tag(bb.device)}:rows{bb.n_rows}:nnz{bb.W_lif.nnz}:rate{bb.n_rate}:D{bb.n_delay}"
            f":ring-{bb.ring_layout}")


def candidate_shapes(fold_modes, steps, grids, n_delay: int, ring_layout: str) -> list:
    """The (fold mode, steps per launch, grid per CU) triples the engine accepts (batch.check_launch_shape)."""
    from .batch import check_launch_shape
    out = []
    for mode, K, gpc in itertools.product(fold_modes, steps, grids):
        try:
            check_launch_shape(mode, K, n_delay, ring_layout)
        except (ValueError, TypeError):
            continue
        out.append({"fold_mode": mode, "steps_per_launch": K, "grid_per_cu": gpc})
    return out


def cached_config(table: dict, key: str, n_delay: int, ring_layout: str):
    """The cached winner for `key`, or None when there is none or the engine would refuse it."""
    entry = table.get(key)
    if entry is None:
        return None
    cfg = entry.get("config", {})
    if not candidate_shapes([cfg.get("fold_mode")], [cfg.get("steps_per_launch")], [cfg.get("grid_per_cu")], n_delay,
                            ring_layout):
        return None
    return entry


def tune_batch(bb, frames: int = 5, fold_modes=None, steps=(4, 5, 6), grids=(64, 96, 128), use_cache=True,
               log=None, warm: int = 30) -> dict:
    """Time each (fold mode, steps per launch, grid per CU) on `bb` (a BatchBrain with rollouts queued) and leave
    it configured with the fastest.  `warm` frames run first: flies start at rest and their firing (the work per
    step) rises over the first tens of frames, so trials compare at a settled activity.  Returns {"config": ...,
    "ms_per_frame": ..., "trials": [...]}.  Raises ValueError when no shape of the search space fits the engine (for
    example every steps-per-launch above its delay)."""
    if fold_modes is None:               # mode 6 needs the slot-major ring; mode 5 is the fastest without it
        fold_modes = (6,) if bb.ring_layout == "slot" else (5,)
    shapes = candidate_shapes(fold_modes, steps, grids, bb.n_delay, bb.ring_layout)
    if not shapes:
        raise ValueError(f"no launch shape to try: fold modes {tuple(fold_modes)}, steps {tuple(steps)} and grids "
                         f"{tuple(grids)} give none the engine accepts (delay {bb.n_delay} steps, "
                         f"{bb.ring_layout}-major ring)")
    key = batch_key(bb)
    table = _load() if use_cache else {}
    hit = cached_config(table, key, bb.n_delay, bb.ring_layout)
    if hit is not None:
        bb.configure(**hit["config"])
        return hit | {"cached": True}
    trials = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if warm:
        bb.run(warm)
    for shape in shapes:
        bb.configure(**shape)
        bb.run(1)                                   # capture + first replay
        torch.cuda.synchronize()
        e0.record()
        bb.run(frames)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / frames
        trials.append(shape | {"ms_per_frame": round(ms, 4)})
        if log:
            log(f"  tune: fold {shape['fold_mode']} K {shape['steps_per_launch']} grid x{shape['grid_per_cu']}: "
                f"{ms:.3f} ms/frame")
    best = min(trials, key=lambda t: t["ms_per_frame"])
    cfg = {k: best[k] for k in ("fold_mode", "steps_per_launch", "grid_per_cu")}
    bb.configure(**cfg)
    result = {"config": cfg, "ms_per_frame": best["ms_per_frame"], "trials": trials}
    if use_cache:
        table[key] = result
        _save(table)
    return result


def packed_key(pb) -> str:
    return f"packed:{_device_tag(pb.device)}:n{pb.n}:nnz{pb.nnz}:k{pb.k}:D{pb.n_delay}"


def tune_packed(pb, steps: int = 180, windows=None, grids=None, use_cache=True, log=None) -> dict:
    """Time each (window, grid) of the packed stepper `pb` (in its current mode) over `steps` steps from rest and
    leave it configured with the fastest (and reset to step 0).  Returns {"config": ..., "ms_per_step": ...,
    "trials": [...]}."""
    key = packed_key(pb)
    table = _load() if use_cache else {}
    if key in table:
        pb.configure(**table[key]["config"])
        return table[key] | {"cached": True}
    cus = torch.cuda.get_device_properties(pb.device).multi_processor_count
    windows = windows or sorted({w for w in (pb.n_delay, (pb.n_delay + 1) // 2) if w >= 1})
    # a wavefront takes one block of 64 cells at a time: grids past one block per wavefront add nothing (grids stay
    # multiples of 8, the XCD count the block schedule is dealt over)
    cap = pb.max_grid if pb.mode == "coop" else max(8, min(cus * 64, -(-pb.words // 32) * 8))
    grids = grids or sorted({min(cap, cus * m) for m in (4, 8, 16, 24, 32, 48, 64)})
    trials = []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for w, g in itertools.product(windows, grids):
        pb.configure(window=w, grid=g)
        pb.run(steps)                               # graph capture + warm caches
        pb.configure(window=w, grid=g)
        e0.record()
        pb.run(steps)
        e1.record()
        torch.cuda.synchronize()
        pb.check()
        ms = e0.elapsed_time(e1) / steps
        trials.append({"window": w, "grid": g, "ms_per_step": round(ms, 5)})
        if log:
            log(f"  tune: window {w} grid {g}: {ms * 1000:.1f} us/step")
    best = min(trials, key=lambda t: t["ms_per_step"])
    cfg = {"window": best["window"], "grid": best["grid"]}
    pb.configure(**cfg)
    result = {"config": cfg, "ms_per_step": best["ms_per_step"], "trials": trials}
    if use_cache:
        table[key] = result
        _save(table)
    return result
