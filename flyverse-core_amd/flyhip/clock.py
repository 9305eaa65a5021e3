"""Frame clocks: how a frame of some ms splits into whole fixed-length steps.

Both the LIF (flyverse/fly.py:421-425) and the optic lobe (flyverse/optic.py:573-575) take the whole steps that fit
in the carried remainder plus the new frame, with a 1e-9 ms allowance for decimal frame lengths such as 0.1, and
carry what is left (never negative) into the next frame.  The step count uses floor(total / dt), not total // dt:
the two differ for decimal steps (1.0 // 0.1 is 9.0).
"""
from __future__ import annotations

import math


def whole_steps(total_ms: float, dt_ms: float) -> int:
    """Whole steps of dt_ms in total_ms."""
    return int(math.floor((total_ms + 1e-9) / dt_ms))


def split_frame(pending_ms: float, frame_ms: float, dt_ms: float) -> tuple[int, float]:
    """(steps, carry) for a frame of frame_ms that follows a carried remainder pending_ms."""
    total = pending_ms + frame_ms
    n = whole_steps(total, dt_ms)
    return n, max(0.0, total - n * dt_ms)
