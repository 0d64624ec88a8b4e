"""Counter-based noise stream shared by the oracle and the HIP step kernel.

TEST INFRASTRUCTURE ONLY: this module is the checker's copy of the noise stream.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it; the product path computes the same stream on the device.

Philox4x32-10 follows Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
easy as 1, 2, 3" (SC'11), Fig. 2 / Table 2: multipliers 0xD2511F53, 0xCD9E8D57,
Weyl key increments 0x9E3779B9, 0xBB67AE85, ten rounds, key bumped before every
round after the first. Pinned by the published known-answer vectors in
``tests/test_oracle_philox.py``.

Stream layout (identical on the device, ``glide_amd/csrc/common.h``):

* key     = (seed & 0xffffffff, seed >> 32)
* counter = (uid, step, stream, 0)
    - ``stream`` = substep index i (0 .. k-1) for the turbulence OU draws,
      ``STREAM_MEANDER`` for the meander draws, ``STREAM_CONVECTION`` for the
      convective redistribution draws.
* uniforms  u_j = ((x_j >> 9) + 0.5) * 2**-23, exact in fp32, in (0, 1).
* normals   Box-Muller on (u0, u1) and (u2, u3):
      n0 = r0 cos(2 pi u1), n1 = r0 sin(2 pi u1), r0 = sqrt(-2 ln u0)
      n2 = r1 cos(2 pi u3), n3 = r1 sin(2 pi u3), r1 = sqrt(-2 ln u2)
* turbulence substep draws (eta_u, eta_v, eta_w) = (n0, n1, n2); this keeps the
  reference's per-substep u, v, w draw order (hanna.py:1335-1353).
* meander draws (eta_um, eta_vm) = (n0, n1) (hanna.py:1040-1051).
* convection draws (u_dest, u_within) = (u0, u1) (emanuel.py:782-801).
"""

from __future__ import annotations

import numpy as np

PHILOX_M0 = np.uint64(0xD2511F53)
PHILOX_M1 = np.uint64(0xCD9E8D57)
PHILOX_W0 = np.uint32(0x9E3779B9)
PHILOX_W1 = np.uint32(0xBB67AE85)

STREAM_MEANDER = 0x10000
STREAM_CONVECTION = 0x20000

_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    """Philox4x32-10 on a batch of counters.

    ``ctr``: uint32 array [..., 4]; ``key``: uint32 array [..., 2] (broadcasts).
    Returns uint32 [..., 4].
    """

    c = np.asarray(ctr, dtype=np.uint32)
    k = np.broadcast_to(np.asarray(key, dtype=np.uint32), c.shape[:-1] + (2,))
    c0 = c[..., 0].astype(np.uint64)
    c1 = c[..., 1].astype(np.uint64)
    c2 = c[..., 2].astype(np.uint64)
    c3 = c[..., 3].astype(np.uint64)
    k0 = k[..., 0].astype(np.uint64)
    k1 = k[..., 1].astype(np.uint64)
    for r in range(10):
        if r > 0:
            k0 = (k0 + np.uint64(PHILOX_W0)) & _MASK32
            k1 = (k1 + np.uint64(PHILOX_W1)) & _MASK32
        p0 = PHILOX_M0 * c0
        p1 = PHILOX_M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0), lo1, (hi0 ^ c3 ^ k1), lo0
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def seed_key(seed: int) -> np.ndarray:
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([s & 0xFFFFFFFF, s >> 32], dtype=np.uint32)


def raw_block(seed: int, uid: np.ndarray, step: int, stream: int) -> np.ndarray:
    """uint32 [N, 4] Philox outputs for counters (uid, step, stream, 0)."""

    uid = np.asarray(uid, dtype=np.uint32)
    ctr = np.zeros(uid.shape + (4,), dtype=np.uint32)
    ctr[..., 0] = uid
    ctr[..., 1] = np.uint32(int(step) & 0xFFFFFFFF)
    ctr[..., 2] = np.uint32(int(stream) & 0xFFFFFFFF)
    return philox4x32_10(ctr, seed_key(seed))


def uniforms(seed: int, uid: np.ndarray, step: int, stream: int) -> np.ndarray:
    """float64 [N, 4] uniforms in (0, 1); each value is exactly representable in fp32."""

    x = raw_block(seed, uid, step, stream)
    return ((x >> np.uint32(9)).astype(np.float64) + 0.5) * (2.0**-23)


def normals(seed: int, uid: np.ndarray, step: int, stream: int) -> np.ndarray:
    """float64 [N, 4] standard normals by Box-Muller on the block's uniforms."""

    u = uniforms(seed, uid, step, stream)
    r0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    r1 = np.sqrt(-2.0 * np.log(u[:, 2]))
    t0 = 2.0 * np.pi * u[:, 1]
    t1 = 2.0 * np.pi * u[:, 3]
    return np.stack([r0 * np.cos(t0), r0 * np.sin(t0), r1 * np.cos(t1), r1 * np.sin(t1)], axis=1)
