"""CPU oracle for the backward-LPDM step path.

TEST INFRASTRUCTURE ONLY. This package is a numpy restatement of the reference's
plain (static-path) method, dtype-parametrised: float64 is the oracle every GPU
result is checked against; float32 is the CPU baseline timed beside the GPU.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import anything under ``oracle/``, and only as the checker. The product
(``glide_amd``) never imports it and fails loudly when its HIP library is missing.

Each function cites the reference file:line it restates. How each is pinned
(known-answer vectors, closed forms, invariants, brute force) is listed in the
test modules ``tests/test_oracle_*.py``.
"""
