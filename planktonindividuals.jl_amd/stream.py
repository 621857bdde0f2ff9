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
KIND_INIT_FE = 19       # IronEnergyMode: the iron quota
KIND_TRACER = 32        # + tracer id; slot = padded cell index
# abiotic particles: species field ABIO_SPECIES + sa; particles_from_bcs! draws
# KIND_ABIO_BC per padded cell (slot = the cell's index)
ABIO_SPECIES = 64
KIND_ABIO_BC = 2

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


# ------------------------------------------------- initial-state transcendentals
# Element-wise mirrors of pi_log_f64 / pi_cos2pi_f64 / pi_init_normal
# (include/pi_math.h): the same binary64 operations in the same order, so the
# host-generated initial state equals the one generated on the device.

def log_f64(x):
    x = np.asarray(x, dtype=np.float64)
    m, k = np.frexp(x)
    lo = m < 0.70710678118654752440
    m = np.where(lo, m + m, m)
    k = np.where(lo, k - 1, k)
    f = m - 1.0
    s = f / (2.0 + f)
    z = s * s
    p = np.full_like(s, 1.0 / 21.0)
    for d in (19.0, 17.0, 15.0, 13.0, 11.0, 9.0, 7.0, 5.0, 3.0):
        p = p * z + 1.0 / d
    p = p * z + 1.0
    kd = k.astype(np.float64)
    return kd * 6.93147180369123816490e-01 + (kd * 1.90821492927058770002e-10 + (s + s) * p)


def cos2pi_f64(b):
    b = np.asarray(b, dtype=np.float64)
    t = 4.0 * b
    q = np.rint(t)
    r = t - q
    x = r * 1.57079632679489661923
    z = x * x
    c = np.full_like(z, 1.0 / 20922789888000.0)
    for v in (-1.0 / 87178291200.0, 1.0 / 479001600.0, -1.0 / 3628800.0, 1.0 / 40320.0, -1.0 / 720.0,
              1.0 / 24.0):
        c = c * z - (-v) if v < 0 else c * z + v
    c = c * z - 0.5
    c = c * z + 1.0
    sn = np.full_like(z, 1.0 / 355687428096000.0)
    for v in (-1.0 / 1307674368000.0, 1.0 / 6227020800.0, -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0,
              1.0 / 120.0, -1.0 / 6.0):
        sn = sn * z - (-v) if v < 0 else sn * z + v
    sn = sn * z + 1.0
    sn = x * sn
    iq = q.astype(np.int64) & 3
    return np.where(iq == 0, c, np.where(iq == 1, -sn, np.where(iq == 2, -c, sn)))


def init_normal(a, b):
    """sqrt(-2 log(1 - a)) cos(2 pi b) in binary64, rounded to Float32."""
    a64 = 1.0 - np.asarray(a, dtype=np.float32).astype(np.float64)
    return (np.sqrt(-2.0 * log_f64(a64)) * cos2pi_f64(np.asarray(b, dtype=np.float32).astype(np.float64))).astype(np.float32)
