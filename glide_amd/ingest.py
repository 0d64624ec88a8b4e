"""Met ingest: source hours decoded on host worker threads into pinned memory,
uploaded on a copy stream, and ordered before their first use on the compute
stream by an event.

Replaces the reference's MetPrefetcher (main.py:791-889), which fetches the
next (backward) hour on one worker thread while the caller works through the
current one, over MetReader.fetch_hourly_window (met_reader.py:438-601). Here:

* ``workers`` threads decode ``lookahead`` hours ahead of the cursor (the
  direction of the walk: backward by default), so a decode slower than the
  GPU's hour (≈0.15 s per 240×160 hour on one core vs ≈10 ms of steps) is
  hidden by parallel decodes;
* a worker copies its arrays into pinned host memory (torch's caching pinned
  allocator, which reuses a block only after the copies that read it have
  completed);
* the caller's thread issues the asynchronous host-to-device copies on one
  copy stream (default priority) and records an event; the compute stream
  waits on that event, so the upload overlaps the steps of the previous hour;
* device hours are kept in an LRU of ``max_resident`` hours; each tensor is
  recorded on the consuming stream, so eviction never frees memory a queued
  kernel still reads.

Results are bit-identical to hours made resident up front (the same decode,
the same bytes). A decode error re-raises on the caller's thread.
"""

from __future__ import annotations

import threading
import time
from collections import OrderedDict
from concurrent.futures import Future, ThreadPoolExecutor

import numpy as np
import torch

from glide_amd.met import convert_source_units

FIELDS = ("src3", "src2", "height", "p_hpa", "conv")
_TORCH_DTYPE = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}


class MetIngest:
    """Streams source hours of a met source (``source.hour(h)`` returning the
    decoded arrays) to the device ahead of a cursor."""

    def __init__(self, source, *, device="cuda", workers: int = 4, lookahead: int = 4, max_resident: int = 8,
                 direction: int = -1, hour_range: tuple[int, int] | None = None):
        if workers < 1 or lookahead < 0 or max_resident < 2:
            raise ValueError("need workers >= 1, lookahead >= 0, max_resident >= 2")
        self.source = source
        self.device = torch.device(device)
        self.lookahead = int(lookahead)
        self.max_resident = int(max_resident)
        self.direction = -1 if direction < 0 else 1
        n_hours = getattr(getattr(source, "spec", None), "n_hours", None)
        self.hour_range = hour_range or (0, int(n_hours) if n_hours is not None else 1 << 30)
        torch.cuda.init()
        self.copy_stream = torch.cuda.Stream(self.device)
        self.pool = ThreadPoolExecutor(max_workers=int(workers), thread_name_prefix="glide-met-ingest")
        self.futures: dict[int, Future] = {}
        self.dev: OrderedDict[int, tuple[dict, torch.cuda.Event]] = OrderedDict()
        self._lock = threading.Lock()
        self.stats = dict(decoded=0, uploads=0, upload_bytes=0, hits=0, wait_s=0.0, decode_s=0.0)

    # -- worker side -------------------------------------------------------
    def _decode(self, h: int) -> dict:
        t0 = time.perf_counter()
        sh = self.source.hour(h)
        arrays = dict(src3=sh.src3, src2=sh.src2, p_hpa=np.asarray(sh.p_level_pa, np.float64) / 100.0,
                      conv=sh.conv)
        if getattr(sh, "units", "final") == "source":  # converted on the device after the upload
            arrays.update(geopotential=sh.geopotential, geopotential_sfc=sh.geopotential_sfc,
                          p_pa=np.asarray(sh.p_level_pa, np.float64))
        else:
            arrays["height"] = sh.height_agl
        pinned = {}
        for k, a in arrays.items():
            t = torch.empty(a.shape, dtype=_TORCH_DTYPE[np.dtype(a.dtype)], pin_memory=True)
            t.numpy()[...] = a
            pinned[k] = t
        with self._lock:
            self.stats["decoded"] += 1
            self.stats["decode_s"] += time.perf_counter() - t0
        return pinned

    def _in_range(self, h: int) -> bool:
        return self.hour_range[0] <= h < self.hour_range[1]

    def prefetch(self, hours) -> None:
        """Start decoding ``hours`` (those not resident or queued)."""

        for h in hours:
            h = int(h)
            if self._in_range(h) and h not in self.dev and h not in self.futures:
                self.futures[h] = self.pool.submit(self._decode, h)

    # -- caller side -------------------------------------------------------
    def get(self, h: int, stream=None) -> dict:
        """Device arrays of source hour ``h`` (src3, src2, height, p_hpa,
        conv), ordered before the work that follows on ``stream`` (a torch
        stream or a raw stream handle; default: the current stream). Starts
        decoding the next ``lookahead`` hours of the walk."""

        h = int(h)
        if not self._in_range(h):
            raise ValueError(f"met hour {h} outside the source's hours {self.hour_range}")
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        elif not isinstance(stream, torch.cuda.Stream):
            stream = torch.cuda.ExternalStream(int(stream), device=self.device) if stream else \
                torch.cuda.default_stream(self.device)
        hit = self.dev.get(h)
        if hit is not None:
            self.dev.move_to_end(h)
            self.stats["hits"] += 1
            arrs, ev = hit
        else:
            fut = self.futures.pop(h, None)
            if fut is None:
                fut = self.pool.submit(self._decode, h)
            t0 = time.perf_counter()
            pinned = fut.result()
            self.stats["wait_s"] += time.perf_counter() - t0
            with torch.cuda.stream(self.copy_stream):
                arrs = {k: torch.empty(v.shape, dtype=v.dtype, device=self.device).copy_(v, non_blocking=True)
                        for k, v in pinned.items()}
                if "geopotential" in arrs:  # a raw hour: the reader's conversions, on the copy stream
                    arrs["height"] = convert_source_units(
                        arrs["src3"], arrs["src2"], arrs.pop("geopotential"), arrs.pop("geopotential_sfc"),
                        arrs.pop("p_pa"), stream=self.copy_stream)
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
            self.stats["uploads"] += 1
            self.stats["upload_bytes"] += sum(v.numel() * v.element_size() for v in pinned.values())
            self.dev[h] = (arrs, ev)
            while len(self.dev) > self.max_resident:
                self.dev.popitem(last=False)
        stream.wait_event(ev)
        for v in arrs.values():
            v.record_stream(stream)
        self.prefetch(h + self.direction * k for k in range(1, self.lookahead + 1))
        return arrs

    def close(self) -> None:
        for f in self.futures.values():
            f.cancel()
        self.futures.clear()
        self.pool.shutdown(wait=True, cancel_futures=True)
        self.dev.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
