/*
 * pi_math.h - numerics every implementation of the hot path must share so that
 * results are reproducible bit for bit between a CPU and the GPU:
 *
 *  1. The shared uniform stream.  The reference draws diffusion and decision
 *     uniforms with `rand!(rng_type(arch), …)` over the whole capacity
 *     (Individuals/Advection/particle_diffusion.jl:18-20,
 *     Plankton/division_death_probability.jl:97-99), and `rng_type(::CPU)`
 *     reseeds from entropy on every call (src/Architectures.jl:33), so no two
 *     runs agree.  Here every draw is a pure function of
 *     (seed, iteration, species, kind, slot): Philox4x32-10 (Salmon, Moraes,
 *     Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) with
 *     counter {slot, iteration, species, kind} and key {seed_lo, seed_hi};
 *     words 0,1,2 of the output are the x/y/z (kind DIFFUSION) or the
 *     graz/mort/dvid (kind DECISION) draws.  A word u becomes a Float32 in
 *     [0,1) the way Julia's rand(Float32) does: (u >> 9 | 0x3f800000) - 1.
 *
 *  2. Transcendentals.  Float32 exp is defined as exp evaluated in binary64
 *     and rounded once to binary32.  The binary64 exp is written here from
 *     its textbook construction (Cody & Waite argument reduction
 *     x = k ln2 + r with a two-word ln2, Taylor polynomial of degree 13 on
 *     |r| <= ln2/2, scaling by 2^k) using only IEEE +, -, *, so every
 *     conforming platform computes identical bits.  `x^4f0` / `x^2f0` follow
 *     Julia's Float32 `^` with an integer-valued exponent: square (and
 *     square again) in binary64, round once.
 *
 *  3. Julia's IEEE semantics for min/max (NaN-propagating, -0 < +0), isless
 *     and `Bool * x` (false * x == copysign(0, x)).
 *
 * Compile every user of this header without FP contraction
 * (-ffp-contract=off) and without fast-math.
 */
#ifndef PI_MATH_H
#define PI_MATH_H

#include <stdint.h>
#include <math.h>

#ifdef __HIPCC__
#define PI_HD __host__ __device__ __forceinline__
#else
#define PI_HD static inline
#endif

/* ---------------------------------------------------------------- stream -- */
enum { PI_KIND_DIFFUSION = 0, PI_KIND_DECISION = 1 };
/* initial-condition sub-streams (iteration field 0): positions, quotas, the two
 * uniforms of the size normal, and per tracer (slot = padded cell index) */
enum { PI_KIND_INIT_POS = 16, PI_KIND_INIT_QUOTA = 17, PI_KIND_INIT_SIZE = 18, PI_KIND_INIT_FE = 19,
       PI_KIND_TRACER = 32 };
/* abiotic particles: species field PI_ABIO_SPECIES + sa; the random walk uses
 * PI_KIND_DIFFUSION, particles_from_bcs! one draw per padded cell (slot = the
 * cell's index), the initial positions PI_KIND_INIT_POS */
enum { PI_ABIO_SPECIES = 64, PI_KIND_ABIO_BC = 2 };
/* particle_release!: the draw of release pair p (species field PI_RELEASE_SPECIES
 * + p, slot = the individual's slot) */
enum { PI_RELEASE_SPECIES = 96, PI_KIND_RELEASE = 3 };

PI_HD void pi_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Julia rand(Float32): 23 random mantissa bits, value in [0, 1). */
PI_HD float pi_u32_to_unit(uint32_t u)
{
    union { uint32_t i; float f; } b;
    b.i = (u >> 9) | 0x3f800000u;
    return b.f - 1.0f;
}

/* Three uniforms of the shared stream. */
PI_HD void pi_uniform3(uint64_t seed, int64_t iteration, int32_t species, int32_t kind,
                       int64_t slot, float* u0, float* u1, float* u2)
{
    uint32_t ctr[4] = { (uint32_t)slot, (uint32_t)iteration, (uint32_t)species, (uint32_t)kind };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t o[4];
    pi_philox4x32_10(ctr, key, o);
    *u0 = pi_u32_to_unit(o[0]);
    *u1 = pi_u32_to_unit(o[1]);
    *u2 = pi_u32_to_unit(o[2]);
}

/* ------------------------------------------------------- transcendentals -- */
PI_HD double pi_exp_f64(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    /* k = nearest integer to x / ln2 */
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01;   /* 32 significant bits: k*ln2_hi exact */
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = rint(x * inv_ln2);
    double r = (x - kd * ln2_hi) - kd * ln2_lo;
    /* Taylor series of e^r, Horner form, degree 13 */
    double p = 1.6059043836821614599e-10;                /* 1/13! */
    p = p * r + 2.0876756987868098979e-09;               /* 1/12! */
    p = p * r + 2.5052108385441718775e-08;               /* 1/11! */
    p = p * r + 2.7557319223985890653e-07;               /* 1/10! */
    p = p * r + 2.7557319223985890653e-06;               /* 1/9!  */
    p = p * r + 2.4801587301587301587e-05;               /* 1/8!  */
    p = p * r + 1.9841269841269841270e-04;               /* 1/7!  */
    p = p * r + 1.3888888888888888889e-03;               /* 1/6!  */
    p = p * r + 8.3333333333333333333e-03;               /* 1/5!  */
    p = p * r + 4.1666666666666666667e-02;               /* 1/4!  */
    p = p * r + 1.6666666666666666667e-01;               /* 1/3!  */
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int k = (int)kd;
    /* 2^k applied in two halves so that 2^k itself never overflows/underflows */
    int k1 = k / 2, k2 = k - k1;
    return ldexp(ldexp(p, k1), k2);
}

PI_HD float pi_expf(float x) { return (float)pi_exp_f64((double)x); }

/* Natural logarithm of a finite x > 0 for the initial-state draws: x = m 2^k
 * with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1), odd
 * series to s^21 (|s| < 0.172: truncation below 1e-17).  Plain binary64
 * operations in a fixed order, so host and device (and the numpy mirror in
 * stream.py) give the same bits; within 2 ulp of the true value. */
PI_HD double pi_log_f64(double x)
{
    int k;
    double m = frexp(x, &k);
    if (m < 0.70710678118654752440) { m = m + m; k = k - 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    p = p * z + 1.0;
    const double kd = (double)k;
    const double ln2_hi = 6.93147180369123816490e-01;   /* k * ln2_hi exact */
    const double ln2_lo = 1.90821492927058770002e-10;
    return kd * ln2_hi + (kd * ln2_lo + (s + s) * p);
}

/* cos(2 pi b) for b in [0, 1): quadrant q = rint(4b), r = 4b - q exact,
 * x = r pi/2 in [-pi/4, pi/4], Taylor series of cos and sin to degree 16/17. */
PI_HD double pi_cos2pi_f64(double b)
{
    const double t = 4.0 * b;
    const double q = rint(t);
    const double r = t - q;
    const double x = r * 1.57079632679489661923;
    const double z = x * x;
    double c = 1.0 / 20922789888000.0;
    c = c * z - 1.0 / 87178291200.0;
    c = c * z + 1.0 / 479001600.0;
    c = c * z - 1.0 / 3628800.0;
    c = c * z + 1.0 / 40320.0;
    c = c * z - 1.0 / 720.0;
    c = c * z + 1.0 / 24.0;
    c = c * z - 0.5;
    c = c * z + 1.0;
    double sn = 1.0 / 355687428096000.0;
    sn = sn * z - 1.0 / 1307674368000.0;
    sn = sn * z + 1.0 / 6227020800.0;
    sn = sn * z - 1.0 / 39916800.0;
    sn = sn * z + 1.0 / 362880.0;
    sn = sn * z - 1.0 / 5040.0;
    sn = sn * z + 1.0 / 120.0;
    sn = sn * z - 1.0 / 6.0;
    sn = sn * z + 1.0;
    sn = x * sn;
    const int iq = ((int)q) & 3;
    return iq == 0 ? c : iq == 1 ? -sn : iq == 2 ? -c : sn;
}

/* The Box-Muller normal of the initial cell size from two stream uniforms
 * a, b in [0, 1): sqrt(-2 log(1 - a)) cos(2 pi b), rounded to Float32. */
PI_HD float pi_init_normal(float a, float b)
{
    const double a64 = 1.0 - (double)a;
    return (float)(sqrt(-2.0 * pi_log_f64(a64)) * pi_cos2pi_f64((double)b));
}

/* Julia Float32 `x ^ 2f0` and `x ^ 4f0` (integer-valued exponent branch). */
PI_HD float pi_pow2f(float x) { double d = (double)x; return (float)(d * d); }
PI_HD float pi_pow4f(float x) { double d = (double)x; double s = d * d; return (float)(s * s); }
PI_HD double pi_pow2d(double x) { return x * x; }
PI_HD double pi_pow4d(double x) { double s = x * x; return s * s; }

/* Float32 x ^ y for a non-integer y (the cell's surface area Sz^(2/3) of
 * IronEnergyMode): exp(y log x) in binary64 with the functions above, rounded
 * once to Float32; 0 for x = 0, +Inf for x = +Inf, NaN for x < 0 or NaN
 * (Julia throws a DomainError for a negative base). */
PI_HD float pi_powf(float x, float y)
{
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return 0.0f;
    if (x == INFINITY) return INFINITY;
    return (float)pi_exp_f64((double)y * pi_log_f64((double)x));
}

/* ------------------------------------------------- Julia IEEE semantics -- */
PI_HD int pi_signbitf(float x) { union { float f; uint32_t i; } b; b.f = x; return (int)(b.i >> 31); }
PI_HD int pi_signbitd(double x) { union { double f; uint64_t i; } b; b.f = x; return (int)(b.i >> 63); }

/* Base.max / Base.min for IEEE floats: NaN wins, max(-0,+0) = +0, min = -0. */
PI_HD float jl_maxf(float x, float y)
{
    float d = x - y;
    if (x != x || y != y) return d;
    return pi_signbitf(d) ? y : x;
}
PI_HD float jl_minf(float x, float y)
{
    float d = x - y;
    if (x != x || y != y) return d;
    return pi_signbitf(d) ? x : y;
}
PI_HD double jl_maxd(double x, double y)
{
    double d = x - y;
    if (x != x || y != y) return d;
    return pi_signbitd(d) ? y : x;
}
PI_HD double jl_mind(double x, double y)
{
    double d = x - y;
    if (x != x || y != y) return d;
    return pi_signbitd(d) ? x : y;
}
/* Base.isless for IEEE floats (total order, NaN last, -0 < +0). */
PI_HD int jl_islessf(float x, float y)
{
    int xn = (x != x), yn = (y != y);
    return (!xn && (yn || (pi_signbitf(x) && !pi_signbitf(y)))) || (x < y);
}
PI_HD int jl_islessd(double x, double y)
{
    int xn = (x != x), yn = (y != y);
    return (!xn && (yn || (pi_signbitd(x) && !pi_signbitd(y)))) || (x < y);
}
/* Base.isequal for IEEE floats: -0 and +0 differ, NaN equals NaN. */
PI_HD int jl_isequalf(float x, float y)
{
    return (x != x && y != y) || (x == y && pi_signbitf(x) == pi_signbitf(y));
}
PI_HD int jl_isequald(double x, double y)
{
    return (x != x && y != y) || (x == y && pi_signbitd(x) == pi_signbitd(y));
}
/* Bool * x  ==  ifelse(b, x, copysign(0, x)) */
PI_HD float jl_maskf(int b, float x) { return b ? x : copysignf(0.0f, x); }
PI_HD double jl_maskd(int b, double x) { return b ? x : copysign(0.0, x); }
/* mod(x, 1) for a float x (Base.mod with a float remainder). */
/* mod(x, 1): rem(x, 1) = x - trunc(x) exactly (the fractional part of a float
 * is a float), then the sign correction of mod */
PI_HD float jl_mod1f(float x)
{
    const float r = x - truncf(x);
    return r == 0.0f ? 0.0f : (r < 0.0f ? r + 1.0f : r);   /* zero: copysign(r, 1) */
}
PI_HD double jl_mod1d(double x)
{
    const double r = x - trunc(x);
    return r == 0.0 ? 0.0 : (r < 0.0 ? r + 1.0 : r);
}

#endif /* PI_MATH_H */
