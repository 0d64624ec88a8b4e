"""GPU: glide_footprint_accumulate_noalloc (workspace in the library's static
device array) against glide_footprint_accumulate (caller workspace), through
the C ABI: the slab, its compensation and the release totals must be bit for
bit equal, over two calls in a row on the same outputs, the second with fewer
pairs (the array keeps the first call's contents), and at the maximum size.
Sizes beyond the maximum are rejected before any launch."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd import _lib as L  # noqa: E402

N_R, CPR = 2, 3 * 48 * 64
N_CELLS = N_R * CPR


def _pairs(n: int, seed: int):
    g = np.random.default_rng(seed)
    keys = g.integers(0, N_CELLS + 1, n).astype(np.uint32)  # N_CELLS: a dropped pair
    keys[: n // 3] = g.integers(0, 40, n // 3)  # crowded cells (a point release)
    vals = (g.random(n, dtype=np.float32) * np.float32(1e-3)).astype(np.float32)
    return (torch.from_numpy(keys.view(np.int32)).cuda(), torch.from_numpy(vals).cuda())


def _outputs():
    return (torch.zeros(N_CELLS, dtype=torch.float32, device="cuda"),
            torch.zeros(N_CELLS, dtype=torch.float32, device="cuda"),
            torch.zeros(N_R, dtype=torch.float64, device="cuda"))


def _with_workspace(keys, vals, out, stream):
    n = keys.numel()
    ws = torch.empty(L.lib().glide_footprint_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    L.check(L.lib().glide_footprint_accumulate(keys.data_ptr(), vals.data_ptr(), n, N_CELLS, CPR, out[0].data_ptr(),
                                               out[1].data_ptr(), out[2].data_ptr(), ws.data_ptr(), ws.numel(),
                                               stream))
    return ws


def _noalloc(keys, vals, out, stream):
    L.check(L.lib().glide_footprint_accumulate_noalloc(keys.data_ptr(), vals.data_ptr(), keys.numel(), N_CELLS, CPR,
                                                       out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                                                       stream))


def test_noalloc_equals_workspace_path_on_two_calls():
    s = torch.cuda.current_stream().cuda_stream
    ref, got = _outputs(), _outputs()
    keep = []
    for n, seed in ((300_000, 1), (70_000, 2)):  # the second call is smaller
        keys, vals = _pairs(n, seed)
        keep.append(_with_workspace(keys, vals, ref, s))
        _noalloc(keys, vals, got, s)
        torch.cuda.synchronize()
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
    assert float(ref[2].sum().item()) > 0.0


def test_noalloc_at_the_maximum_size():
    s = torch.cuda.current_stream().cuda_stream
    n = L.GLIDE_NOALLOC_MAX_PAIRS
    keys, vals = _pairs(n, 3)
    ref, got = _outputs(), _outputs()
    _with_workspace(keys, vals, ref, s)
    _noalloc(keys, vals, got, s)
    torch.cuda.synchronize()
    for a, b in zip(ref, got):
        assert torch.equal(a, b)


def test_noalloc_rejects_more_pairs_than_its_workspace():
    lib = L.lib()
    rc = lib.glide_footprint_accumulate_noalloc(None, None, L.GLIDE_NOALLOC_MAX_PAIRS + 1, N_CELLS, CPR, None, None,
                                                None, None)
    assert rc == L.GLIDE_ERR_INVALID
    assert b"maximum" in lib.glide_last_error()
