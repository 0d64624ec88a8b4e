"""CPU: the per-release scale of the fixed-point footprint. For each release
r with residence bound T_r = (sum of its weights) * (longest backward length +
2 dt), 2^e_r must put T_r in (2^60, 2^61]: the whole release fits int64 with
a factor 4 to spare, and the scale is no coarser than that (an exponent one
too small or too large, or a floor for the ceiling, fails). Releases without
particles get 0."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from glide_amd.runner import BatchEngine


def test_fixed_exponents_put_the_bound_in_range():
    rng = np.random.default_rng(3)
    n_r = 7
    rel = rng.integers(0, n_r - 1, size=5000)  # release n_r - 1 has no particles
    w = rng.uniform(1e-6, 3.0, size=rel.size).astype(np.float32)
    fake = SimpleNamespace(grid_cfg={"n_r": n_r}, ra=SimpleNamespace(rel_idx=rel.astype(np.int32), weight=w),
                           length_s=86400, cfg={"dt": 60})
    e = BatchEngine.fixed_exponents(fake)
    assert e.dtype == np.int32 and e.shape == (n_r,)
    for r in range(n_r):
        t = float(np.abs(w[rel == r].astype(np.float64)).sum()) * (86400 + 120)
        if t == 0:
            assert e[r] == 0
            continue
        scaled = t * 2.0 ** int(e[r])
        assert 2.0**60 < scaled <= 2.0**61, (r, t, e[r], scaled)


def test_fixed_exponent_of_a_unit_release():
    # 2^20 particles of weight 2^-20 over 24 h at dt = 60 s: T = 86520 s,
    # log2 T = 16.4, e = 61 - 17 = 44, and one particle-step (v = 60 * 2^-20)
    # becomes 60 * 2^24, an exact integer far above the rounding unit
    n = 1 << 20
    fake = SimpleNamespace(grid_cfg={"n_r": 1},
                           ra=SimpleNamespace(rel_idx=np.zeros(n, np.int32), weight=np.full(n, 2.0**-20, np.float32)),
                           length_s=86400, cfg={"dt": 60})
    e = BatchEngine.fixed_exponents(fake)
    assert e.tolist() == [44]
    assert np.float64(np.float32(60 * 2.0**-20)) * 2.0**44 == 60 * 2**24

