"""GPU: batches of releases with different particle counts and backward
lengths in one engine (the network batch), against the float64 oracle, against
each release run alone, and with the hourly compaction of finished particles.

Bars: footprint per release within 1 % relative L2 of the oracle; a release
alone and in a batch stepped with per-release termini (noise keyed by the
release's particle ids and absolute time): every particle bit-identical,
footprints within float32 summation-order rounding (2e-5); compaction on/off:
the same. Residence totals: float32 and float64 trajectories of the same noise
drift apart, and at 8k-16k particles per release, released between 10 and 300
m (some of them straddle the 100 m top of the output layer), the totals then
differ by up to 5e-4 (measured; mixed signs), while the single 10k-particle
release of the plain run sits at 1e-5 against its 1e-4 bar. Bars here: 1e-3
per release, and 1e-4 on the mean signed error over the releases (no bias;
measured 1.2e-5).

Every engine test runs with both footprint accumulations: "fixed" (exact
fixed-point sums inside the step kernel) and "sort" (radix sort plus
fixed-order float reduction); the overlap and binned-key options belong to the
sort accumulation."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd.runner import BatchEngine  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402
from oracle import run as ORUN  # noqa: E402

STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "alive")
TOTAL_BAR = 1e-3
BIAS_BAR = 1e-4


@pytest.fixture(scope="module")
def cfg_met():
    cfg = workload_config("batched-cpu-6r")
    return cfg, SyntheticMet(cfg["spec"])


@pytest.fixture(scope="module", params=["fixed", "sort"])
def batch(request, cfg_met):
    cfg, met = cfg_met
    eng = BatchEngine(cfg, device="cuda", met=met, compact=True, accumulate=request.param)
    eng.run()
    torch.cuda.synchronize()
    return cfg, met, eng


def _mode(eng) -> str:
    return "fixed" if eng.fixed_mode else "sort"


@pytest.fixture(scope="module")
def oracle_batch(cfg_met):
    cfg, met = cfg_met
    return ORUN.run(cfg, dty=np.float64, met=met)


def _slabs(eng, cfg):
    g = cfg["grid"]
    return eng.footprint.cpu().numpy().astype(np.float64).reshape(g["n_r"], -1)


def test_network_batch_matches_oracle(batch, oracle_batch):
    cfg, met, eng = batch
    lens = {r.length_s for r in cfg["releases"]}
    assert len(lens) > 1, "the batch must mix backward lengths"
    ref = oracle_batch
    fp = _slabs(eng, cfg)
    fr = ref.footprint.reshape(fp.shape)
    tot = eng.release_total.cpu().numpy()
    l2 = [np.linalg.norm(fp[r] - fr[r]) / np.linalg.norm(fr[r]) for r in range(fp.shape[0])]
    dt = [(tot[r] - fr[r].sum()) / fr[r].sum() for r in range(fp.shape[0])]
    print("per-release footprint L2", np.round(l2, 5), "total rel. error", np.round(dt, 7))
    for r in range(fp.shape[0]):
        assert fr[r].sum() > 0, r
        assert l2[r] <= 0.01, (r, l2[r])
        assert abs(dt[r]) <= TOTAL_BAR, (r, tot[r], fr[r].sum())
    assert abs(np.mean(dt)) <= BIAS_BAR, dt
    assert int(eng.escaped.cpu().numpy().sum()) == int(sum(ref.escaped_per_step))
    # the live set shrank as releases finished
    assert eng.n_live < eng.n


def test_release_alone_equals_release_in_batch(batch):
    """A release run alone and inside a batch stepped with each release's own
    terminus (release_terminus=True): same noise (particle ids and absolute
    time), same met, same steps. Every particle is bit-identical; footprints
    differ only in float32 summation order inside a cell. (Under the
    reference's batch terminus a release other than the earliest-ending one is
    stepped once more, main.py:1383.)"""
    cfg, met, b = batch
    eng = BatchEngine(cfg, device="cuda", met=met, compact=True, release_terminus=True, accumulate=_mode(b))
    eng.run()
    S = eng.state_numpy()
    fp = _slabs(eng, cfg)
    tot = eng.release_total.cpu().numpy()
    for k in (0, 4):
        rel = cfg["releases"][k]
        one = dataclasses.replace(rel, release_idx=0, noise_id=rel.release_idx)
        g1 = dict(cfg["grid"], n_r=1)
        c1 = dict(cfg, releases=[one], grid=g1)
        e1 = BatchEngine(c1, device="cuda", met=met, accumulate=_mode(b))
        e1.run()
        s1 = e1.state_numpy()
        sel = np.isin(S["uid"], s1["uid"])
        assert sel.sum() == one.n_particles
        for name in STATE:
            assert np.array_equal(S[name][sel], s1[name]), (k, name)
        f1 = e1.footprint.cpu().numpy().astype(np.float64)
        assert np.array_equal(f1 != 0, fp[k] != 0), k
        np.testing.assert_allclose(f1, fp[k], rtol=2e-5, atol=0)
        t1 = float(e1.release_total.cpu().numpy()[0])
        assert abs(t1 - tot[k]) <= 1e-9 * tot[k], (k, t1, tot[k])


def test_compaction_leaves_results_unchanged(batch):
    cfg, met, eng = batch
    plain = BatchEngine(cfg, device="cuda", met=met, compact=False, accumulate=_mode(eng))
    plain.run()
    a, b = plain.state_numpy(), eng.state_numpy()
    for name in STATE:
        assert np.array_equal(a[name], b[name]), name
    assert np.array_equal(plain.escaped.cpu().numpy(), eng.escaped.cpu().numpy())
    fa, fb = plain.footprint.cpu().numpy(), eng.footprint.cpu().numpy()
    assert np.array_equal(fa != 0, fb != 0)
    np.testing.assert_allclose(fb, fa, rtol=2e-5, atol=0)


def test_network_batch_bf16_met_within_bar(batch, oracle_batch):
    """The precision layout of the batched method (bfloat16 wind and support
    records, compensated float32 footprint, float64 totals) against the float64
    oracle on float32 met: footprint within 1 % relative L2 per release (the
    batched method's bar); totals reported (bfloat16 rounding of the winds is a
    small systematic change of the transport, not summation noise)."""
    cfg, met, b = batch
    from glide_amd import _lib as L

    eng = BatchEngine(cfg, device="cuda", met=met, met_dtype=L.GLIDE_MET_BF16, compensated=True, compact=True,
                      accumulate=_mode(b))
    eng.run()
    ref = oracle_batch
    fp = _slabs(eng, cfg)
    if eng.comp is not None:  # Neumaier terms of the sort accumulation
        fp = fp + eng.comp.cpu().numpy().astype(np.float64).reshape(cfg["grid"]["n_r"], -1)
    fr = ref.footprint.reshape(fp.shape)
    tot = eng.release_total.cpu().numpy()
    l2 = [np.linalg.norm(fp[r] - fr[r]) / np.linalg.norm(fr[r]) for r in range(fp.shape[0])]
    dt = [(tot[r] - fr[r].sum()) / fr[r].sum() for r in range(fp.shape[0])]
    print("bf16 met: per-release footprint L2", np.round(l2, 5), "total rel. error", np.round(dt, 6))
    for r in range(fp.shape[0]):
        assert l2[r] <= 0.01, (r, l2[r])  # the batched method's bar
        assert abs(dt[r]) <= 0.01, (r, dt[r])


@pytest.mark.parametrize("use_graph", [False, True])
def test_overlapped_accumulation_is_bit_identical(batch, use_graph):
    """overlap=True runs each step's footprint accumulation on a second stream,
    concurrently with the next step: the same particles, footprint, totals and
    counts, bit for bit (each accumulation still sums its own step's keys in
    the same fixed order; only the streams differ)."""
    cfg, met, eng = batch
    if eng.fixed_mode:
        pytest.skip("overlap is an option of the sort accumulation")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ov = BatchEngine(cfg, device="cuda", met=met, compact=True, overlap=True, use_graph=use_graph,
                         accumulate="sort")
        ov.run()
    torch.cuda.synchronize()
    a, b = eng.state_numpy(), ov.state_numpy()
    for name in STATE:
        assert np.array_equal(a[name], b[name]), name
    assert np.array_equal(eng.footprint.cpu().numpy(), ov.footprint.cpu().numpy())
    assert np.array_equal(eng.release_total.cpu().numpy(), ov.release_total.cpu().numpy())
    assert np.array_equal(eng.escaped.cpu().numpy(), ov.escaped.cpu().numpy())
    assert np.array_equal(eng.active_steps.cpu().numpy(), ov.active_steps.cpu().numpy())
    ov.close()


def test_binned_keys_equal_flat_keys(batch):
    """Release-binned footprint keys (time bin in a per-release table) sort in
    the same order as the flat cell keys, so the footprint, compensation and
    totals are bit-identical to the flat-key run."""
    cfg, met, eng = batch
    if eng.fixed_mode:
        pytest.skip("binned keys are an option of the sort accumulation")
    flat = BatchEngine(cfg, device="cuda", met=met, compact=True, binned_keys=False, accumulate="sort")
    flat.run()
    torch.cuda.synchronize()
    assert eng.binned_keys and not flat.binned_keys
    assert np.array_equal(eng.footprint.cpu().numpy(), flat.footprint.cpu().numpy())
    assert np.array_equal(eng.release_total.cpu().numpy(), flat.release_total.cpu().numpy())
    a, b = eng.state_numpy(), flat.state_numpy()
    for name in STATE:
        assert np.array_equal(a[name], b[name]), name
