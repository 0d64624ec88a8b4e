"""GPU: the fixed-point footprint accumulation (GLIDE_F_FIXED), bit for bit
against an exact integer restatement on the CPU.

A sort-accumulating engine with flat keys and a fixed-point engine step the
same particles (the accumulation does not touch them, so both hold the same
particles after every step). After each step the sort engine's keys and values
are that step's contributions. The CPU turns each value v of release r into
rint(v * 2^e_r) -- exact in float64, since v is a float32 and the scale a power
of two -- and adds the integers per cell with np.add.at on int64 (exact).
The device's int64 cells must equal these after every step; the float
footprint must equal numpy's correctly rounded int64 -> float32 conversion
scaled by 2^-e_r, and the release totals the exact per-release integer sums in
float64. The steps include the hourly convection, whose kernel emits the
step's residence on those steps, the hourly compaction and re-binning.

Then the invariances the integer sums give: launch shape, re-binning interval,
the number of accumulator copies and graph replay leave the fixed-point
footprint bit-identical, and the sort accumulation's float32 sums agree with
it within float32 summation rounding.
"""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd.runner import BatchEngine, ConvectionOptions  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402


@pytest.fixture(scope="module")
def setup():
    cfg = dict(workload_config("batched-cpu-6r"), convection=True)
    return cfg, SyntheticMet(cfg["spec"])


def _exact_contributions(keys, vals, e, cpr, n_cells):
    ok = keys < n_cells
    k = keys[ok].astype(np.int64)
    v = vals[ok].astype(np.float64) * np.exp2(e[k // cpr].astype(np.float64))
    return k, np.rint(v).astype(np.int64)


def test_fixed_cells_equal_exact_integer_sums(setup):
    cfg, met = setup
    kw = dict(device="cuda", met=met, compact=True, reorder_every=8, convection=ConvectionOptions())
    srt = BatchEngine(cfg, accumulate="sort", binned_keys=False, **kw)
    fix = BatchEngine(cfg, accumulate="fixed", device_met=srt.met, **kw)
    e = fix.fixed_exp.cpu().numpy()
    cpr, n_cells = fix.cells_per_release, fix.n_cells
    acc = np.zeros(n_cells, np.int64)
    n_contrib = 0
    for i in range(len(fix.sched)):
        srt.run_step()
        fix.run_step()
        torch.cuda.synchronize()
        n = srt.n_live
        k, q = _exact_contributions(srt.keys[:n].cpu().numpy().view(np.uint32), srt.vals[:n].cpu().numpy(), e,
                                    cpr, n_cells)
        np.add.at(acc, k, q)
        n_contrib += k.size
        if i % 10 == 0 or i == len(fix.sched) - 1:
            assert np.array_equal(fix.fixed_cells().cpu().numpy(), acc), i
    assert n_contrib > 0 and acc.max() > 2**24  # integers well beyond float32's exact range
    # float footprint: one correctly rounded conversion per cell, then the exact 2^-e scale
    want = np.ldexp(acc.astype(np.float32), -np.repeat(e, cpr)).astype(np.float32)
    got = fix.footprint.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # totals: exact per-release integer sums, converted once to float64
    rel = acc.reshape(-1, cpr).sum(axis=1)
    want_t = np.array([np.ldexp(float(t), -int(x)) for t, x in zip(rel, e)])
    assert np.array_equal(fix.release_total.cpu().numpy(), want_t)
    # the particles the two engines stepped are the same
    a, b = srt.state_numpy(), fix.state_numpy()
    for name in ("lon", "lat", "z", "alive"):
        assert np.array_equal(a[name], b[name]), name
    # the sort accumulation's float32 sums (per step, then over steps) agree
    # within float32 summation rounding
    fs = srt.footprint.cpu().numpy()
    assert np.array_equal(fs != 0, got != 0)
    np.testing.assert_allclose(fs, got, rtol=2e-5, atol=0)


@pytest.mark.parametrize("n", [1, 63, 65, 1000])
def test_fixed_cells_ragged_particle_counts(n):
    """Particle counts that leave a wavefront partly empty (1, 63, 65) or
    several whole: the lanes past n and the inactive particles add nothing,
    and the cells equal the exact integer sums of the sort path's keys and
    values, step by step (first 40 steps of the 3 h plain workload)."""
    import dataclasses

    base = workload_config("plain-cpu-10k-3h")
    rel = dataclasses.replace(base["releases"][0], n_particles=n)
    cfg = dict(base, releases=[rel])
    met = SyntheticMet(cfg["spec"])
    srt = BatchEngine(cfg, device="cuda", met=met, accumulate="sort", binned_keys=False)
    fix = BatchEngine(cfg, device="cuda", met=met, device_met=srt.met)
    e = fix.fixed_exp.cpu().numpy()
    acc = np.zeros(fix.n_cells, np.int64)
    for _ in range(min(40, len(fix.sched))):
        srt.run_step()
        fix.run_step()
        torch.cuda.synchronize()
        k, q = _exact_contributions(srt.keys[:n].cpu().numpy().view(np.uint32), srt.vals[:n].cpu().numpy(), e,
                                    fix.cells_per_release, fix.n_cells)
        np.add.at(acc, k, q)
        assert np.array_equal(fix.fixed_cells().cpu().numpy(), acc)
    assert acc.sum() > 0


def test_fixed_footprint_invariant_to_launch_and_order(setup):
    cfg, met = setup
    kw = dict(device="cuda", met=met, compact=True, convection=ConvectionOptions())
    ref = BatchEngine(cfg, **kw)
    ref.run()
    base = ref.fixed_cells().cpu().numpy()
    fp = ref.footprint.cpu().numpy()
    tot = ref.release_total.cpu().numpy()
    s = torch.cuda.Stream()
    for extra in (dict(step_block=64, fixed_copies=1), dict(step_block=512, reorder_every=4, fixed_copies=3),
                  dict(use_graph=True, reorder_every=16), dict(reorder_every=8, bin_by_release=False)):
        with torch.cuda.stream(s):
            eng = BatchEngine(cfg, device_met=ref.met, **kw, **extra)
            eng.run()
        torch.cuda.synchronize()
        assert np.array_equal(eng.fixed_cells().cpu().numpy(), base), extra
        assert np.array_equal(eng.footprint.cpu().numpy(), fp), extra
        assert np.array_equal(eng.release_total.cpu().numpy(), tot), extra
        if extra.get("use_graph"):
            assert eng.graph_launches > 0
        eng.close()


def test_fixed_exponents_bound_the_release_residence(setup):
    cfg, met = setup
    eng = BatchEngine(cfg, device="cuda", met=met, compact=True)
    eng.run()
    e = eng.fixed_exp.cpu().numpy()
    cells = eng.fixed_cells().cpu().numpy().reshape(cfg["grid"]["n_r"], -1)
    for r in range(cells.shape[0]):
        tot = int(cells[r].astype(object).sum())  # Python integers: no wrap
        assert 0 <= tot < 2**61, (r, tot)
        assert tot > 2**40, (r, tot)  # the scale leaves the sums far above the rounding unit
    assert np.all(e > 0)


def test_fixed_finalize_rounds_correctly_through_the_abi():
    """glide_fixed_finalize on chosen int64 cells: the copies are summed
    exactly, each cell is rounded once to the nearest float32 (ties to even:
    2^25 + 1 -> 2^25, 2^25 + 3 -> 2^25 + 4) and scaled by 2^-e of its
    release, and the totals are the exact integer sums in float64. Checked
    bit for bit against numpy's correctly rounded int64 conversions."""
    from glide_amd import _lib as L

    rng = np.random.default_rng(5)
    n_r, cpr, copies = 3, 4099, 3
    n_cells = n_r * cpr
    # a release's cells sum below 2^62 (the exponent rule's promise), so the
    # integer totals cannot wrap; the chosen values reach 2^61
    cells = rng.integers(-(2**48), 2**48, size=(copies, n_cells), dtype=np.int64)
    ties = np.array([2**25 + 1, 2**25 + 3, 2**40 + 2**16, -(2**25 + 1), 0, 1, 2**61 - 1], np.int64)
    cells[:, : ties.size] = 0
    cells[0, : ties.size] = ties
    e = np.array([44, -3, 61], np.int32)
    fixed = torch.from_numpy(cells.reshape(-1)).cuda()
    exp = torch.from_numpy(e).cuda()
    slab = torch.empty(n_cells, dtype=torch.float32, device="cuda")
    tot = torch.empty(n_r, dtype=torch.float64, device="cuda")
    ws = torch.empty(max(8, L.lib().glide_fixed_workspace_bytes(n_r)), dtype=torch.uint8, device="cuda")
    L.check(L.lib().glide_fixed_finalize(fixed.data_ptr(), copies, n_cells, cpr, exp.data_ptr(), slab.data_ptr(),
                                         tot.data_ptr(), ws.data_ptr(), ws.numel(), L.stream_ptr()))
    torch.cuda.synchronize()
    s = cells.sum(axis=0)  # |values| < 2^62: no wrap
    assert all(abs(sum(int(v) for v in s[r * cpr:(r + 1) * cpr])) < 2**62 for r in range(n_r))
    want = np.ldexp(s.astype(np.float32), -np.repeat(e, cpr)).astype(np.float32)
    assert np.float32(2**25 + 1) == np.float32(2**25) and np.float32(2**25 + 3) == np.float32(2**25 + 4)
    assert np.array_equal(slab.cpu().numpy().view(np.uint32), want.view(np.uint32))
    want_t = [np.ldexp(float(int(x)), -int(k)) for x, k in zip([sum(int(v) for v in s[r * cpr:(r + 1) * cpr])
                                                                for r in range(n_r)], e)]
    assert np.array_equal(tot.cpu().numpy(), np.array(want_t))
    # argument checks: copies out of range, cells_per_release not dividing
    with pytest.raises(ValueError):
        L.check(L.lib().glide_fixed_finalize(fixed.data_ptr(), 2000, n_cells, cpr, exp.data_ptr(), slab.data_ptr(),
                                             None, None, 0, L.stream_ptr()))
    with pytest.raises(ValueError):
        L.check(L.lib().glide_fixed_finalize(fixed.data_ptr(), copies, n_cells, cpr + 1, exp.data_ptr(),
                                             slab.data_ptr(), None, None, 0, L.stream_ptr()))
