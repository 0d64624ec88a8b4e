"""CPU: the autotuner's choice among timed trials: the fastest, except that
trials within 1 % of it go to the one that re-bins least often."""

from __future__ import annotations


def test_autotune_tie_rule_prefers_less_rebinning():
    from glide_amd.autotune import _choose

    t = [dict(step_block=256, reorder_every=32, particle_steps_per_s=100.0),
         dict(step_block=256, reorder_every=0, particle_steps_per_s=99.5),
         dict(step_block=128, reorder_every=8, particle_steps_per_s=99.8)]
    assert _choose(t)["reorder_every"] == 0  # within 1 %: the least re-binning
    t[1]["particle_steps_per_s"] = 90.0
    assert _choose(t) == t[0]  # re-binning every 32 steps beats every 8, and 0 is 10 % slower
