"""Vectorised host implementation of the shared uniform stream of include/pi_math.h
(Philox4x32-10, counter {slot, iteration, species, kind}, key {seed_lo, seed_hi},
Julia rand(Float32) mapping).  Used to draw initial conditions on the host and
to give callers the exact draws the device consumes for a given step.
"""
import numpy as np

KIND_DIFFUSION = 0
KIND_DECISION = 1
# initial-condition sub-streams (host only): iteration field = 0, kinds >= 16
KIND_INIT_POS = 16      # x, y, z
KIND_INIT_QUOTA = 17    # Cq, Nq, Pq
KIND_INIT_SIZE = 18     # two uniforms for the Box-Muller normal of Sz
KIND_TRACER = 32        # + tracer id; slot = padded cell index

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint32(0x9E3779B9)
_W1 = np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 on uint32 arrays (broadcasting); returns four uint32 arrays."""
    c0, c1, c2, c3 = (np.asarray(a, dtype=np.uint32) for a in (c0, c1, c2, c3))
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    c0, c1, c2, c3 = (a.copy() for a in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for r in range(10):
            if r > 0:
                k0 = np.uint32((int(k0) + int(_W0)) & 0xFFFFFFFF)
                k1 = np.uint32((int(k1) + int(_W1)) & 0xFFFFFFFF)
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
            lo0 = (p0 & _MASK).astype(np.uint32)
            hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
            lo1 = (p1 & _MASK).astype(np.uint32)
            n0 = hi1 ^ c1 ^ k0
            n2 = hi0 ^ c3 ^ k1
            c0, c1, c2, c3 = n0, lo1, n2, lo0
    return c0, c1, c2, c3


def to_unit(u):
    """Julia rand(Float32): (u >> 9 | 0x3f800000) reinterpreted, minus 1."""
    b = (np.asarray(u, dtype=np.uint32) >> np.uint32(9)) | np.uint32(0x3F800000)
    return b.view(np.float32) - np.float32(1.0)


def uniform3(seed, iteration, species, kind, slots):
    """The three Float32 uniforms of the shared stream for each slot in `slots`.
    Large requests are split over threads (numpy releases the GIL in ufuncs)."""
    slots = np.asarray(slots, dtype=np.uint64)
    if slots.size >= (1 << 21):
        import os
        from concurrent.futures import ThreadPoolExecutor
        nth = max(1, min(16, os.cpu_count() or 1))
        parts = np.array_split(slots, nth * 4)
        with ThreadPoolExecutor(nth) as ex:
            res = list(ex.map(lambda p: _uniform3(seed, iteration, species, kind, p), parts))
        return tuple(np.concatenate([r[i] for r in res]) for i in range(3))
    return _uniform3(seed, iteration, species, kind, slots)


def _uniform3(seed, iteration, species, kind, slots):
    k0 = np.uint32(seed & 0xFFFFFFFF)
    k1 = np.uint32((seed >> 32) & 0xFFFFFFFF)
    o0, o1, o2, _ = philox4x32_10((slots & _MASK).astype(np.uint32), np.uint32(iteration & 0xFFFFFFFF),
                                  np.uint32(species), np.uint32(kind), k0, k1)
    return to_unit(o0), to_unit(o1), to_unit(o2)
