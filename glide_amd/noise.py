"""Noise keys of the drop-in schemes' calls (HannaScheme.step,
EmanuelReducedConvection.maybe_convect).

The engine's Philox stream is keyed by (seed, particle id) and counted by
(step, stream); the drop-in interfaces (turbulence/base.py:60-78,
convection/base.py:59-96) pass no step number, only the met window (its start
and end times) and the bracket weight ``t_alpha``. The step counter is
therefore model time, floor(t / dt) of the time t = t_start + t_alpha (t_end -
t_start) at which the weight is evaluated (the step midpoint, main.py:892-938),
so a call's draws do not depend on how many calls came before it.

Two call patterns would otherwise reuse draws, and each moves the key's high
32 bits instead: a call at the same model time as the previous one (a caller
stepping a fixed window, as the reference's well-mixed tests do) counts a
repeat, and a call later in time than the previous one (a new backward sweep,
such as the next release batch, whose particles reuse the row ids) counts a
sweep. A single backward integration draws with both counts 0.
"""

from __future__ import annotations

import math

from glide_amd.met import MetFieldMetadata


def model_time_seconds(md: MetFieldMetadata, t_alpha: float) -> float:
    t0 = md.time_start.timestamp()
    t1 = md.time_end.timestamp()
    return t0 + float(t_alpha) * (t1 - t0)


def model_time_step(md: MetFieldMetadata, t_alpha: float, dt_seconds: float) -> int:
    """The Philox step counter of a call at ``t_alpha`` in window ``md``:
    floor(t / dt), t in seconds since the Unix epoch, modulo 2^31."""

    return int(math.floor(model_time_seconds(md, t_alpha) / float(dt_seconds))) % (1 << 31)


class NoiseClock:
    """Step counter and key of one scheme instance's successive calls."""

    def __init__(self) -> None:
        self.last_t: float | None = None
        self.sweep = 0
        self.repeat = 0

    def next(self, md: MetFieldMetadata, t_alpha: float, dt_seconds: float, seed: int) -> tuple[int, int]:
        """(step counter, Philox key) of the next call."""

        t = model_time_seconds(md, t_alpha)
        tol = 1e-3 * float(dt_seconds)
        if self.last_t is not None:
            if t > self.last_t + tol:  # later than the previous call: a new backward sweep
                self.sweep += 1
                self.repeat = 0
            elif t >= self.last_t - tol:  # the same model time again
                self.repeat += 1
            else:
                self.repeat = 0
        self.last_t = t
        hi = ((self.sweep & 0xFFFF) << 16) | (self.repeat & 0xFFFF)
        key = (int(seed) & 0xFFFFFFFFFFFFFFFF) ^ (hi << 32)
        return model_time_step(md, t_alpha, dt_seconds), key
