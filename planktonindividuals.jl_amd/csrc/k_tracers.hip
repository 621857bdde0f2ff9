/*
 * k_tracers.hip - tracer_update! (Biogeochemistry/tracer_update.jl:22-43) on gfx950.
 *
 *   k_sweep_x   DST3FL x sweep (calc_Gcsˣ! + multi_dim_x!, Advection/multi_dim_adv.jl:125-190)
 *               for all 14 tracers.  Lanes own FACES: a row of Nx cells has
 *               Nx + 1 faces, a wave takes 64 consecutive faces and updates the
 *               63 cells whose west face it holds (the east face comes from the
 *               next lane by a shuffle), so every face flux is computed once.
 *   k_sweep_y   y sweep; each lane walks a run of cells along y, computing each
 *               face flux once and carrying it (and the stencil) to the next cell.
 *   k_fill_line halo fill of one dimension along lines, as fill_halo_*
 *               (HaloRegions/halo_regions.jl) with broadcast unaliasing (each
 *               line is owned by one lane, which reads its sources first).
 *   k_xy        fused x and y sweeps (grids with halo 2, Nx <= 256, no aliased
 *               periodic direction), rows staged in LDS, one level per block row.
 *   k_final     z sweep + Gcs = temp - c + diffusion + sinking + forcing +
 *               apply_tendency, four lanes per column run (tracer groups closed
 *               under the forcings) walking along z with the advective and
 *               diffusive z fluxes and the z stencil carried, writing the next
 *               tracers and, on k_xy's grids, their halo images.
 *
 * Exactness: every value is the reference's expression in its operation order.
 *  - The face flux ½(T+|T|)(…) + ½(T−|T|)(…) is evaluated for the sign of T
 *    only.  For 0 < T < 2^126, ½(T+|T|) = T exactly and ½(T−|T|) = +0.  With the
 *    four stencil values below 2^100 in magnitude and the Courant number below
 *    2^60, every Ψ is a number in [0, 1] (no inf − inf, no 0·inf), so the
 *    second factor is finite and its product a signed zero: the sum equals the
 *    first product whenever that is non-zero.  Symmetrically for T < 0.  Every
 *    other case (T = ±0, a zero product, large or non-finite values) takes the
 *    full formula.
 *  - Divisions by fixed divisors (Δx, Δy, dz, the volume, 6, and C + 1e-20
 *    shared by the 14 tracers of a face) use divc (pi_device.h).
 *  - The face flux at the last interior face is computed from the halo values
 *    in Periodic directions (identical to the reference's Gc halo copy of the
 *    first face) and is zero in Bounded ones (fill_halo_east_Gc!, halo_regions.jl:22-24).
 */
#include "pi_device.h"
#include "k_args.h"


/* Tracer buffers are slabs of 14 padded fields (field q at slab + q·ncell) and
 * are addressed as a uniform base plus a 32-bit byte offset (ncell·4 < 2^32 is
 * checked when the model is created): the loads take the SGPR-base form and
 * no 64-bit address is kept per tracer. */
__device__ __forceinline__ const float* fld(const float* slab, const DevGrid& g, int q) { return slab + (long long)q * g.ncell; }
__device__ __forceinline__ float* fld(float* slab, const DevGrid& g, int q) { return slab + (long long)q * g.ncell; }
__device__ __forceinline__ float ld(const float* base, uint32_t cell) { return *(const float*)((const char*)base + cell * 4u); }
__device__ __forceinline__ void st(float* base, uint32_t cell, float v) { *(float*)((char*)base + cell * 4u) = v; }

/* Every kernel here is instantiated twice: LAND = false (no land cell, the
 * masks are identities and each face is straight-line code) and LAND = true. */
template <bool LAND>
__device__ __forceinline__ float lmask(const DevGrid& g, uint32_t outer, uint32_t inner, float x)
{
    if (!LAND) return x;
    return jl_maskf(g.lm[outer] != 0, jl_maskf(g.lm[inner] != 0, x));
}

/* face quantities of DST3FL.jl:27-38, shared by the 14 tracers */
struct Face {
    float C, d0, d1, T, omc, Ce;   /* Courant number, δ⁰, δ¹, transport, 1 − C, C + 1e-20 */
    double yCe;                    /* RN(1 / (C + 1e-20)) */
    bool fast;                     /* T ≠ ±0, |T| < 2^126, C < 2^60 */
};

__device__ __forceinline__ Face face(const DevGrid& g, float u, float dt, float D, double yD, float A)
{
    Face f;
    f.C = fabsf(divc(u * dt, D, yD));
    f.d0 = divc((2.0f - f.C) * (1.0f - f.C), 6.0f, g.y6);
    f.d1 = divc(1.0f - f.C * f.C, 6.0f, g.y6);
    f.T = A * u;
    f.omc = 1.0f - f.C;
    f.Ce = f.C + 1.0e-20f;
    f.yCe = 1.0 / (double)f.Ce;
    f.fast = f.C < 0x1p60f && ((f.T > 0.0f && f.T < 0x1p126f) || (f.T < 0.0f && f.T > -0x1p126f));
    return f;
}

/* the same with the Courant number's spacing D from a metric array (LatLonGrid):
 * an IEEE division */
__device__ __forceinline__ Face face_gm(const DevGrid& g, float u, float dt, float D, float A)
{
    Face f;
    f.C = fabsf((u * dt) / D);
    f.d0 = divc((2.0f - f.C) * (1.0f - f.C), 6.0f, g.y6);
    f.d1 = divc(1.0f - f.C * f.C, 6.0f, g.y6);
    f.T = A * u;
    f.omc = 1.0f - f.C;
    f.Ce = f.C + 1.0e-20f;
    f.yCe = 1.0 / (double)f.Ce;
    f.fast = f.C < 0x1p60f && ((f.T > 0.0f && f.T < 0x1p126f) || (f.T < 0.0f && f.T > -0x1p126f));
    return f;
}

/* θ and Ψ (DST3FL.jl:23-57) */
__device__ __forceinline__ float theta(float dn, float dd)
{
    const float tmax = 1.0e20f;
    const float q = dn / dd;                       /* both arms evaluated: no branch */
    const float cap = copysignf(tmax, dn * dd);
    return fabsf(dd) * tmax <= fabsf(dn) ? cap : q;
}
template <bool EXACT>
__device__ __forceinline__ float psi(const Face& F, float th, bool& sub)
{
    return dmaxf(0.0f, dminf(dminf(1.0f, F.d0 + F.d1 * th), divq<EXACT>(th * F.omc, F.Ce, F.yCe, sub)));
}

__device__ __forceinline__ bool small(float x) { return fabsf(x) < 0x1p100f; }   /* false for NaN */

/* The flux of one tracer at one face is
 *   ½(T+|T|)(bp + Ψ(θ(dnp, dd0))·dd0) + ½(T−|T|)(bm − Ψ(θ(dnm, dd0))·dd0)
 *   x/y (DST3FL.jl:85-91): bp = c₋₁, dnp = δ₋₁, bm = c₀,  dnm = δ₊₁
 *   z   (DST3FL.jl:93-95): bp = c₀,  dnp = δ₊₁, bm = c₋₁, dnm = δ₋₁ */
struct Sten { float dd0, bp, dnp, bm, dnm; bool ok; };

template <bool LAND, bool Z>
__device__ __forceinline__ Sten stencil_of(const DevGrid& g, uint32_t f, uint32_t s, float cm2, float cm1, float c0,
                                           float cp1)
{
    Sten r;
    if (Z) {
        r.dd0 = lmask<LAND>(g, f - s, f, cm1 - c0);
        r.dnm = lmask<LAND>(g, f - 2 * s, f - s, cm2 - cm1);
        r.dnp = lmask<LAND>(g, f, f + s, c0 - cp1);
        r.bp = c0; r.bm = cm1;
    } else {
        r.dd0 = lmask<LAND>(g, f - s, f, c0 - cm1);
        r.dnp = lmask<LAND>(g, f - 2 * s, f - s, cm1 - cm2);
        r.dnm = lmask<LAND>(g, f, f + s, cp1 - c0);
        r.bp = cm1; r.bm = c0;
    }
    r.ok = small(cm2) && small(cm1) && small(c0) && small(cp1);
    return r;
}
/* the stencil of face f of field p: the values at f − 2s, f − s, f, f + s */
template <bool LAND, bool Z>
__device__ __forceinline__ Sten stencil(const DevGrid& g, const float* p, uint32_t f, uint32_t s)
{
    return stencil_of<LAND, Z>(g, f, s, ld(p, f - 2 * s), ld(p, f - s), ld(p, f), ld(p, f + s));
}
/* the term of the sign of T only, with divq; bad = the full formula (with
 * exact division) is needed */
__device__ __forceinline__ float flux_fast(const Face& F, const Sten& s, bool& bad)
{
    const bool pos = F.T > 0.0f;
    bool sub = false;
    const float p = psi<false>(F, theta(pos ? s.dnp : s.dnm, s.dd0), sub);
    const float pd = p * s.dd0;
    const float r = F.T * (pos ? s.bp + pd : s.bm - pd);
    bad = !(s.ok && F.fast && r != 0.0f) || sub;
    return r;
}
__device__ __forceinline__ float flux_full(const Face& F, const Sten& s)
{
    const float T = F.T;
    bool sub = false;
    const float pp = psi<true>(F, theta(s.dnp, s.dd0), sub), pm = psi<true>(F, theta(s.dnm, s.dd0), sub);
    return 0.5f * (T + fabsf(T)) * (s.bp + pp * s.dd0) + 0.5f * (T - fabsf(T)) * (s.bm - pm * s.dd0);
}
/* the fluxes of tracers Q0..Q1-1 of slab at face f (between f − s and f);
 * zero = Bounded end face.  The full formula is evaluated only where needed,
 * after the fast terms of all these tracers, so the face is one block of
 * straight-line code. */
template <bool LAND, bool Z, int Q0, int Q1>
__device__ __forceinline__ void face_fluxes(const DevGrid& g, const Face& F, bool zero, uint32_t f, uint32_t s,
                                            const float* slab, float* out)
{
    uint32_t need = 0;
#pragma unroll
    for (int q = Q0; q < Q1; ++q) {
        const Sten t = stencil<LAND, Z>(g, fld(slab, g, q), f, s);
        bool bad;
        const float r = flux_fast(F, t, bad);
        out[q - Q0] = zero ? 0.0f : r;
        need |= (uint32_t)(bad && !zero) << q;
    }
    if (__builtin_expect(need != 0, 0)) {
#pragma unroll
        for (int q = Q0; q < Q1; ++q)
            if ((need >> q) & 1u) out[q - Q0] = flux_full(F, stencil<LAND, Z>(g, fld(slab, g, q), f, s));
    }
}

/* ----------------------------------------------------------------- x sweep */
template <bool LAND, int Q0, int Q1>
__device__ __forceinline__ void sweep_x_half(const SweepArgs& a, const DevGrid& g, const Face& Fc, bool zero, bool upd,
                                             uint32_t c, uint32_t cf, float coef, float dT, const float* C0)
{
    float fl[Q1 - Q0];
    face_fluxes<LAND, false, Q0, Q1>(g, Fc, zero, cf, 1, C0, fl);
#pragma unroll
    for (int q = Q0; q < Q1; ++q) {
        const float f1 = __shfl_down(fl[q - Q0], 1, 64);
        const float c0 = ld(fld(C0, g, q), c);
        if (upd) st(fld(a.out, g, q), c, c0 - coef * (lmask<LAND>(g, c, c + 1, f1 - fl[q - Q0]) - c0 * dT));
    }
}

template <bool LAND, bool GM>
__device__ __forceinline__ void k_sweep_x_body(const SweepArgs& a, Blk K)
{
    const DevState* S = a.st;
    if (S->err) return;
    const DevGrid& g = a.g;
    const long long nrow = (long long)g.Ny * g.Nz;
    const long long nface = nrow * (g.Nx + 1);
    const long long gt = K.b * (long long)blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const long long F = (gt >> 6) * 63 + lane;
    const bool valid = F < nface;
    const long long Fv = valid ? F : nface - 1;
    const long long row = Fv / (g.Nx + 1);
    const int fi = (int)(Fv - row * (g.Nx + 1));          /* 0..Nx: west face of interior cell fi */
    const int i = fi + 1 + g.Hx;
    const int j = (int)(row % g.Ny) + 1 + g.Hy;
    const int k = (int)(row / g.Ny) + 1 + g.Hz;
    const uint32_t c = (uint32_t)gidx(g, i, j, k);
    /* Periodic with Nx < Hx: the last face takes the first face's flux, as the
     * reference's Gc halo copy does; its halo stencil would read halo planes
     * that the aliased fill leaves holding earlier values. */
    const bool wrap = g.tx == PI_PERIODIC && g.Nx < g.Hx && fi == g.Nx;
    const uint32_t cf = wrap ? c - (uint32_t)g.Nx : c;
    const float* C0 = S->tr[S->tcur][0];
    const float dt = a.dt;
    /* Trans_x = Ax·u, CFLx = |u·ΔT/ΔxC| at the face (DST3FL.jl:4,13) */
    const Face Fc = GM ? face_gm(g, ld(a.vel, cf), dt, g.mdxC[col2(g, wrap ? i - g.Nx : i, j)], g.mAx[cf])
                       : face(g, ld(a.vel, cf), dt, g.dx, g.ydx, g.dy * dzF_at(g, k));
    const bool zero = fi == g.Nx && g.tx == PI_BOUNDED;
    const bool upd = valid && lane < 63 && fi < g.Nx;
    const float T1 = __shfl_down(Fc.T, 1, 64);
    const float dT = lmask<LAND>(g, c, c + 1, T1 - Fc.T);
    const float coef = GM ? dt / g.mVol[c] : divc(dt, g.vol[k - 1], g.yvol[k - 1]);
    sweep_x_half<LAND, 0, 7>(a, g, Fc, zero, upd, c, cf, coef, dT, C0);
    sweep_x_half<LAND, 7, 14>(a, g, Fc, zero, upd, c, cf, coef, dT, C0);
}
__device__ __forceinline__ void k_sweep_x0_body(const SweepArgs& a, Blk K) { k_sweep_x_body<false, false>(a, K); }
__device__ __forceinline__ void k_sweep_x1_body(const SweepArgs& a, Blk K) { k_sweep_x_body<true, false>(a, K); }
__device__ __forceinline__ void k_sweep_x2_body(const SweepArgs& a, Blk K) { k_sweep_x_body<false, true>(a, K); }
__device__ __forceinline__ void k_sweep_x3_body(const SweepArgs& a, Blk K) { k_sweep_x_body<true, true>(a, K); }
PI_KERNEL2(k_sweep_x0, k_sweep_x0_body, SweepArgs, 256)
PI_KERNEL2(k_sweep_x1, k_sweep_x1_body, SweepArgs, 256)
/* LatLonGrid: areas, spacings and volumes from the metric arrays */
PI_KERNEL2(k_sweep_x2, k_sweep_x2_body, SweepArgs, 256)
PI_KERNEL2(k_sweep_x3, k_sweep_x3_body, SweepArgs, 256)

/* ----------------------------------------------------------------- y sweep */
#define RUN_Y 16

template <bool LAND, int Q0, int Q1>
__device__ __forceinline__ void sweep_y_half(const SweepArgs& a, const DevGrid& g, const Face& Fb, const Face& Fa,
                                             bool zero, uint32_t c, uint32_t cf, uint32_t sy, float coef,
                                             const float* C0, float* fl)
{
    float f[Q1 - Q0];
    face_fluxes<LAND, false, Q0, Q1>(g, Fb, zero, cf, sy, a.prev, f);
    const uint32_t cs = c - sy;    /* the cell south of face c */
    const float dT = lmask<LAND>(g, cs, c, Fb.T - Fa.T);
#pragma unroll
    for (int q = Q0; q < Q1; ++q) {
        st(fld(a.out, g, q), cs,
           ld(fld(a.prev, g, q), cs) - coef * (lmask<LAND>(g, cs, c, f[q - Q0] - fl[q]) - ld(fld(C0, g, q), cs) * dT));
        fl[q] = f[q - Q0];
    }
}

template <bool LAND, bool GM>
__device__ __forceinline__ void k_sweep_y_body(const SweepArgs& a, Blk K)
{
    const DevState* S = a.st;
    if (S->err) return;
    const DevGrid& g = a.g;
    const int nrun = (g.Ny + RUN_Y - 1) / RUN_Y;
    const long long nthr = (long long)g.Nx * nrun * g.Nz;
    const uint32_t sy = (uint32_t)g.Px;
    const float* C0 = S->tr[S->tcur][0];
    const float dt = a.dt;
    const bool wrapy = g.ty == PI_PERIODIC && g.Ny < g.Hy;    /* the last face copies the first, as in k_sweep_x */
    const uint32_t back = (uint32_t)g.Ny * sy;
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < nthr; t += (long long)K.n * blockDim.x) {
        const int i = (int)(t % g.Nx) + 1 + g.Hx;
        const int r = (int)((t / g.Nx) % nrun);
        const int k = (int)(t / ((long long)g.Nx * nrun)) + 1 + g.Hz;
        const int j0 = g.Hy + 1 + r * RUN_Y;
        const int j1 = min(j0 + RUN_Y, g.Ny + g.Hy + 1);     /* cells j0..j1-1, faces j0..j1 */
        const float A = g.dx * dzF_at(g, k);
        const float coef = GM ? 0.0f : divc(dt, g.vol[k - 1], g.yvol[k - 1]);
        uint32_t c = (uint32_t)gidx(g, i, j0, k);            /* cell north of the current face */
        float fl[PI_NTRACERS];                               /* fluxes of the previous face */
        /* Trans_y = Ay·v, CFLy = |v·ΔT/ΔyC| at the face (DST3FL.jl:5,14) */
        Face Fa = GM ? face_gm(g, ld(a.vel, c), dt, g.mdyC[col2(g, i, j0)], g.mAy[c])
                     : face(g, ld(a.vel, c), dt, g.dy, g.ydy, A);
        face_fluxes<LAND, false, 0, 7>(g, Fa, false, c, sy, a.prev, fl);
        face_fluxes<LAND, false, 7, 14>(g, Fa, false, c, sy, a.prev, fl + 7);
#pragma unroll 1
        for (int jj = j0 + 1; jj <= j1; ++jj) {
            c += sy;
            const bool wj = wrapy && jj == g.Ny + g.Hy + 1;
            const uint32_t cf = wj ? c - back : c;
            const Face Fb = GM ? face_gm(g, ld(a.vel, cf), dt, g.mdyC[col2(g, i, wj ? jj - g.Ny : jj)], g.mAy[cf])
                               : face(g, ld(a.vel, cf), dt, g.dy, g.ydy, A);
            const bool zero = jj == g.Ny + g.Hy + 1 && g.ty == PI_BOUNDED;
            const float cf_coef = GM ? dt / g.mVol[c - sy] : coef;     /* the cell south of the face */
            sweep_y_half<LAND, 0, 7>(a, g, Fb, Fa, zero, c, cf, sy, cf_coef, C0, fl);
            sweep_y_half<LAND, 7, 14>(a, g, Fb, Fa, zero, c, cf, sy, cf_coef, C0, fl);
            Fa = Fb;
        }
    }
}
__device__ __forceinline__ void k_sweep_y0_body(const SweepArgs& a, Blk K) { k_sweep_y_body<false, false>(a, K); }
__device__ __forceinline__ void k_sweep_y1_body(const SweepArgs& a, Blk K) { k_sweep_y_body<true, false>(a, K); }
__device__ __forceinline__ void k_sweep_y2_body(const SweepArgs& a, Blk K) { k_sweep_y_body<false, true>(a, K); }
__device__ __forceinline__ void k_sweep_y3_body(const SweepArgs& a, Blk K) { k_sweep_y_body<true, true>(a, K); }
PI_KERNEL2(k_sweep_y0, k_sweep_y0_body, SweepArgs, 256)
PI_KERNEL2(k_sweep_y1, k_sweep_y1_body, SweepArgs, 256)
PI_KERNEL2(k_sweep_y2, k_sweep_y2_body, SweepArgs, 256)
PI_KERNEL2(k_sweep_y3, k_sweep_y3_body, SweepArgs, 256)

/* fill_halo_{west,east} / {south,north} / {top,bottom} of one dimension, on the
 * lines given by (o1, o2) ranges.  lines_full = 1 covers every padded line (the
 * reference fills whole planes); 0 covers interior lines only. */
__device__ __forceinline__ void k_fill_line_body(const FillArgs& a, Blk K)
{
    const DevState* S = a.st;
    if (S->err) return;
    const DevGrid& g = a.g;
    const int N = a.dim == 0 ? g.Nx : a.dim == 1 ? g.Ny : g.Nz;
    const int H = a.dim == 0 ? g.Hx : a.dim == 1 ? g.Hy : g.Hz;
    const int topo = a.dim == 0 ? g.tx : a.dim == 1 ? g.ty : g.tz;
    const long long stride = a.dim == 0 ? 1 : a.dim == 1 ? g.Px : (long long)g.Px * g.Py;
    const int P1 = a.dim == 0 ? g.Py : g.Px, P2 = a.dim == 2 ? g.Py : g.Pz;
    const int N1 = a.dim == 0 ? g.Ny : g.Nx, N2 = a.dim == 2 ? g.Ny : g.Nz;
    const int H1 = a.dim == 0 ? g.Hy : g.Hx, H2 = a.dim == 2 ? g.Hy : g.Hz;
    const long long n1 = a.lines_full ? P1 : N1, n2 = a.lines_full ? P2 : N2;
    const long long nlines = n1 * n2 * a.nf;
    float* slab = a.f ? a.f : S->tr[S->tcur ^ 1][0];
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < nlines; t += (long long)K.n * blockDim.x) {
        const int q = (int)(t / (n1 * n2));
        const long long l = t % (n1 * n2);
        const int o1 = (int)(l % n1) + 1 + (a.lines_full ? 0 : H1);
        const int o2 = (int)(l / n1) + 1 + (a.lines_full ? 0 : H2);
        long long base;    /* index of position 1 along the line */
        if (a.dim == 0) base = gidx(g, 1, o1, o2);
        else if (a.dim == 1) base = gidx(g, o1, 1, o2);
        else base = gidx(g, o1, o2, 1);
        float* c = fld(slab, g, q);
        #define AT(p) c[base + (long long)((p) - 1) * stride]
        if (topo == PI_PERIODIC) {
            float tmp[8];
            for (int h = 0; h < H; ++h) tmp[h] = AT(N + 1 + h);          /* west <- c[N+1:N+H] */
            for (int h = 0; h < H; ++h) AT(1 + h) = tmp[h];
            for (int h = 0; h < H; ++h) tmp[h] = AT(H + 1 + h);          /* east <- c[1+H:2H] */
            for (int h = 0; h < H; ++h) AT(N + H + 1 + h) = tmp[h];
        } else {
            const float lo = AT(H + 1);
            for (int h = 0; h < H; ++h) AT(1 + h) = lo;
            const float hi = AT(N + H);
            for (int h = 0; h < H; ++h) AT(N + H + 1 + h) = hi;
        }
        #undef AT
    }
}
PI_KERNEL(k_fill_line, FillArgs, 256)

/* Copies every halo cell of the current tracers into slab f (null = the tracers
 * written by this step).  Used only where a Periodic direction has N < H: there
 * the reference's halo fill reads halo planes of the array it fills, which hold
 * the values of its previous fill (tracer_temp is a copy of the tracers, and
 * apply_tendency! writes the interior only), so the slab to be filled starts
 * from the current tracers' halos. */
__device__ __forceinline__ void k_halo_copy_body(const FillArgs& a, Blk K)
{
    const DevState* S = a.st;
    if (S->err) return;
    const DevGrid& g = a.g;
    const float* src = S->tr[S->tcur][0];
    float* dst = a.f ? a.f : S->tr[S->tcur ^ 1][0];
    const long long n = (long long)g.ncell * a.nf;
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)K.n * blockDim.x) {
        const long long cell = t % g.ncell;
        const int i = (int)(cell % g.Px), j = (int)((cell / g.Px) % g.Py), k = (int)(cell / ((long long)g.Px * g.Py));
        const bool inside = i >= g.Hx && i < g.Hx + g.Nx && j >= g.Hy && j < g.Hy + g.Ny && k >= g.Hz && k < g.Hz + g.Nz;
        if (!inside) dst[t] = src[t];
    }
}
PI_KERNEL(k_halo_copy, FillArgs, 256)

/* fill_halo_tracer! of the new tracers (HaloRegions/halo_regions.jl:37-47: x,
 * then y, then z fills over whole planes) for halo 2 and Periodic N >= 2: the
 * composition leaves in every halo cell the value of the interior cell whose
 * index is, per direction, the periodic image (Periodic) or the nearest edge
 * cell (Bounded).  halo_src is that source index per direction (0-based padded
 * coordinates); k_final writes each interior value to its halo images (the
 * inverse map, halo_images). */
__device__ __forceinline__ int halo_src(int p, int H, int N, int topo)
{
    int x = p - H;                                                /* 0-based interior coordinate, may be outside */
    if (topo == PI_PERIODIC) x = x < 0 ? x + N : x >= N ? x - N : x;
    else x = x < 0 ? 0 : x >= N ? N - 1 : x;
    return x + H;
}

/* ---------------------------------------------- z sweep + the rest of the update */

/* the tracers that receive particle exchange (plk), tracer_update.jl:22-43;
 * DFe and PFe_bio also in IronEnergyMode models (the FE variants of k_final) */
__device__ __forceinline__ constexpr bool has_plk(int q)
{
    return q == PI_DIC || q == PI_NH4 || q == PI_NO3 || q == PI_PO4 || q == PI_DOC || q == PI_DON ||
           q == PI_DOP || q == PI_POC || q == PI_PON || q == PI_POP;
}

struct ZCell {      /* cell cc (level k; column i, j; all 1-based padded) above face c */
    uint32_t cc, c;
    int k, i, j;
    int it;         /* the model's iteration during this step (flux boundary conditions) */
    float dT, coef, khAx, khAy, inv_vol, dzf, vol;
    double ydzf, yvol;
    /* LatLonGrid: κv·Az of the column, κh·Ax / κh·Ay of the east and north
     * faces (khAx, khAy: west and south), ΔxF / ΔyF of the four faces */
    float kvAz, khAxE, khAyN, dxFW, dxFE, dyFS, dyFN;
};

/* x / volume(i, j, k) of the cell of z: by its reciprocal (divc) on a
 * RectilinearGrid, an IEEE division on a LatLonGrid */
template <bool GM>
__device__ __forceinline__ float per_vol(float x, const ZCell& z) { return GM ? x / z.vol : divc(x, z.vol, z.yvol); }

/* getbc (HaloRegions/boundary_conditions.jl:31-40) of tracer q on face f at
 * face position (p, r) (1-based interior, face array size np x nr) */
__device__ __forceinline__ float bc_at(const BcTab* T, int q, int f, int p, int r, int np, int nr, int it)
{
    const int kind = T->kind[q][f];
    if (kind == PI_BCK_CONST) return T->val[q][f];
    const long long t = kind == PI_BCK_SERIES ? (long long)(it - 1) : 0ll;
    return T->arr[q][f][(long long)(p - 1) + (long long)np * ((long long)(r - 1) + (long long)nr * t)];
}

/* apply_bcs! (HaloRegions/apply_bcs.jl:5-47, 87-99) for tracer q at the cell of
 * z: in the order west, east, south, north, bottom, top, Gc += or -=
 * ((bc·ΔT)·A)/V at the interior cells next to the face; only faces with a
 * boundary condition (bit f of a.bcm[q]) are looked at */
template <bool GM>
__device__ __forceinline__ float flux_bcs(const FinalArgs& a, const DevGrid& g, int q, const ZCell& z, float G)
{
    const BcTab* T = a.bc;
    const uint32_t m = a.bcm[q];
    const int i = z.i - g.Hx, j = z.j - g.Hy, k = z.k - g.Hz;
    const float dt = a.dt;
    const float Ax = GM ? g.mAx[z.cc] : g.dy * dzF_at(g, z.k), Ay = GM ? g.mAy[z.cc] : g.dx * dzF_at(g, z.k);
    const float Az = GM ? g.mAz[col2(g, z.i, z.j)] : g.dx * g.dy;
    if ((m & (1u << PI_BC_WEST)) && i == 1)
        G = G + per_vol<GM>((bc_at(T, q, PI_BC_WEST, j, k, g.Ny, g.Nz, z.it) * dt) * Ax, z);
    if ((m & (1u << PI_BC_EAST)) && i == g.Nx)
        G = G - per_vol<GM>((bc_at(T, q, PI_BC_EAST, j, k, g.Ny, g.Nz, z.it) * dt) * Ax, z);
    if ((m & (1u << PI_BC_SOUTH)) && j == 1)
        G = G + per_vol<GM>((bc_at(T, q, PI_BC_SOUTH, i, k, g.Nx, g.Nz, z.it) * dt) * Ay, z);
    if ((m & (1u << PI_BC_NORTH)) && j == g.Ny)
        G = G - per_vol<GM>((bc_at(T, q, PI_BC_NORTH, i, k, g.Nx, g.Nz, z.it) * dt) * Ay, z);
    if ((m & (1u << PI_BC_BOTTOM)) && k == g.Nz)
        G = G + per_vol<GM>((bc_at(T, q, PI_BC_BOTTOM, i, j, g.Nx, g.Ny, z.it) * dt) * Az, z);
    if ((m & (1u << PI_BC_TOP)) && k == 1)
        G = G - per_vol<GM>((bc_at(T, q, PI_BC_TOP, i, j, g.Nx, g.Ny, z.it) * dt) * Az, z);
    return G;
}

/* Gcs, diffusion, sinking, forcing and apply_tendency (and the diffusive flux
 * of the face below) of tracer q at the cell above face z.c, from its operands:
 * cq, cn = c at the cell and below it, xm/xp/ym/yp = its horizontal neighbours,
 * pv = the z-sweep input at the cell, fzq/fq = advective fluxes of its top and
 * bottom faces, dzq = diffusive flux of its top face, csq = c above the cell
 * (sinking tracers), pk = plk. */
template <bool LAND, bool BCS, bool GM>
__device__ __forceinline__ void cell_update(const FinalArgs& a, const DevGrid& g, const ZCell& z, int q, const float* tt,
                                            float cq, float cn, float xm, float xp, float ym, float yp, float pv,
                                            float fzq, float fq, float dzq, float csq, float pk,
                                            float& out, float& d)
{
    const uint32_t cc = z.cc, c = z.c, sy = (uint32_t)g.Px;
    const int k = z.k;
    const float dt = a.dt;
    const float kvAz = GM ? z.kvAz : a.kv * (g.dx * g.dy);
    /* κv·Az·(δz c / dzF) at face c (Advection/tracer_diffusion.jl:1-30) */
    d = kvAz * divc(lmask<LAND>(g, cc, c, cq - cn), z.dzf, z.ydzf);
    const float temp3 = pv - z.coef * (lmask<LAND>(g, cc, c, fzq - fq) - cq * z.dT);
    /* Gcs = tracer_temp - tracer (calc_tracer_tendency!) */
    float G = temp3 - cq;
    float fxa, fxb, fya, fyb;
    if (GM) {       /* κ·A(i)·(δ⁰c(i) / ΔF(i)) at the west/east and south/north faces */
        fxa = z.khAx * (lmask<LAND>(g, cc - 1, cc, cq - xm) / z.dxFW);
        fxb = z.khAxE * (lmask<LAND>(g, cc, cc + 1, xp - cq) / z.dxFE);
        fya = z.khAy * (lmask<LAND>(g, cc - sy, cc, cq - ym) / z.dyFS);
        fyb = z.khAyN * (lmask<LAND>(g, cc, cc + sy, yp - cq) / z.dyFN);
    } else {
        fxa = z.khAx * divc(lmask<LAND>(g, cc - 1, cc, cq - xm), g.dx, g.ydx);
        fxb = z.khAx * divc(lmask<LAND>(g, cc, cc + 1, xp - cq), g.dx, g.ydx);
        fya = z.khAy * divc(lmask<LAND>(g, cc - sy, cc, cq - ym), g.dy, g.ydy);
        fyb = z.khAy * divc(lmask<LAND>(g, cc, cc + sy, yp - cq), g.dy, g.ydy);
    }
    const float ddx = lmask<LAND>(g, cc, cc + 1, fxb - fxa);
    const float ddy = lmask<LAND>(g, cc, cc + sy, fyb - fya);
    const float ddz = lmask<LAND>(g, cc, c, dzq - d);
    const float lap = z.inv_vol * (ddx + ddy + ddz);
    G = G + lap * dt;
    /* tracer_sinking! (Advection/tracer_sinking.jl:1-43) */
    if (q == PI_PFe_inorg || q == PI_Dust || q == PI_PFe_bio) {
        const float ws = q == PI_PFe_bio ? a.wsink_org : a.wsink_inorg;
        const float f0 = (k - 1 <= g.Hz) ? 0.0f : ws * csq;
        const float f1 = (k >= g.Nz + g.Hz) ? 0.0f : ws * cq;
        G = G + div_dzC(g, f0 - f1, k - g.Hz) * dt;
    }
    /* tracer_forcing! (tracer_forcings.jl:2-50) */
    switch (q) {
        case PI_DIC: G = G + tt[PI_DOC] * a.kDOC * dt; break;
        case PI_DOC: G = G - tt[PI_DOC] * a.kDOC * dt + tt[PI_POC] * a.kPOC * dt; break;
        case PI_POC: G = G - tt[PI_POC] * a.kPOC * dt; break;
        case PI_NH4: G = G + tt[PI_DON] * a.kDON * dt - tt[PI_NH4] * a.Nit * dt; break;
        case PI_NO3: G = G + tt[PI_NH4] * a.Nit * dt; break;
        case PI_DON: G = G - tt[PI_DON] * a.kDON * dt + tt[PI_PON] * a.kPON * dt; break;
        case PI_PON: G = G - tt[PI_PON] * a.kPON * dt; break;
        case PI_PO4: G = G + tt[PI_DOP] * a.kDOP * dt; break;
        case PI_DOP: G = G - tt[PI_DOP] * a.kDOP * dt + tt[PI_POP] * a.kPOP * dt; break;
        case PI_POP: G = G - tt[PI_POP] * a.kPOP * dt; break;
        case PI_DFe: {
            const float DFe = tt[PI_DFe];
            const float ex = dmaxf((DFe - a.lig), 0.0f);
            G = G + tt[PI_PFe_bio] * a.kPOC * dt -
                (a.lPOC * tt[PI_POC] + a.lmin + a.ldust * tt[PI_Dust]) * DFe * a.DFeFrac * dt +
                tt[PI_PFe_inorg] * a.kdiss * dt -
                ex * DFe * a.DFeFrac * a.lFe * dt;
            break;
        }
        case PI_PFe_bio:
            G = G + tt[PI_POC] * tt[PI_DFe] * a.DFeFrac * a.lPOC * dt - a.kPOC * tt[PI_PFe_bio] * dt;
            break;
        case PI_PFe_inorg: {
            const float DFe = tt[PI_DFe];
            const float ex = dmaxf((DFe - a.lig), 0.0f);
            G = G - tt[PI_PFe_inorg] * a.kdiss * dt +
                (a.lmin + a.ldust * tt[PI_Dust]) * DFe * a.DFeFrac * dt +
                ex * DFe * a.DFeFrac * a.lFe * dt;
            break;
        }
        default: break;
    }
    /* apply_bcs! (tracer_update.jl:37-38): boundary cells of tracers with a flux
     * boundary condition */
    if (BCS) {
        const bool edge = z.i == g.Hx + 1 || z.i == g.Hx + g.Nx || z.j == g.Hy + 1 || z.j == g.Hy + g.Ny ||
                          z.k == g.Hz + 1 || z.k == g.Hz + g.Nz;
        if (edge && a.bcm[q]) G = flux_bcs<GM>(a, g, q, z, G);
    }
    /* apply_tendency_kernel! (tracer_update.jl:6-20); the tracers without
     * particle exchange add a zero plk */
    out = cq + (G + per_vol<GM>(pk, z));
}

__device__ __forceinline__ int sink_index(int q) { return q == PI_PFe_inorg ? 0 : q == PI_Dust ? 1 : q == PI_PFe_bio ? 2 : -1; }

/* The halo cells whose fill_halo_tracer! value is a copy of interior
 * coordinate s (0-based padded) in one direction, as a mask over the
 * candidates k = 0: s itself, 1, 2: halo cells 0, 1, 3, 4: halo cells N+2,
 * N+3 (halo 2): the inverse of halo_src. */
__device__ __forceinline__ int halo_cand(int k, int s, int H, int N) { return k == 0 ? s : k < 3 ? k - 1 : N + H + (k - 3); }
__device__ __forceinline__ uint32_t halo_images(int s, int H, int N, int topo)
{
    uint32_t m = 1u;
#pragma unroll
    for (int k = 1; k < 5; ++k) m |= (uint32_t)(halo_src(halo_cand(k, s, H, N), H, N, topo) == s) << k;
    return m;
}
struct HaloImg { uint32_t mx, my; int sx, sy; };   /* masks and the column (0-based padded) */

/* k_final runs four threads per column run, one per tracer group.  The
 * forcings (tracer_forcings.jl:2-50) couple tracers only within these groups:
 * DIC DOC POC; NH4 NO3 DON PON; PO4 DOP POP; and DFe PFe_inorg PFe_bio Dust,
 * whose forcings also read POC.  fgrp(G, i) = the i-th tracer of group G,
 * fg_n(G) its size. */
#define FG_MAX 4
#define FG_NGRP 4
__device__ __forceinline__ constexpr int fg_n(int G) { return G == 0 || G == 2 ? 3 : 4; }
__device__ __forceinline__ constexpr int fgrp(int G, int i)
{
    return G == 0 ? (i == 0 ? PI_DIC : i == 1 ? PI_DOC : PI_POC)
         : G == 1 ? (i == 0 ? PI_NH4 : i == 1 ? PI_NO3 : i == 2 ? PI_DON : PI_PON)
         : G == 2 ? (i == 0 ? PI_PO4 : i == 1 ? PI_DOP : PI_POP)
                  : (i == 0 ? PI_DFe : i == 1 ? PI_PFe_inorg : i == 2 ? PI_PFe_bio : PI_Dust);
}

/* face_fluxes (vertical) for the tracers of group G at face f from the values
 * v[i][0..3] of tracer fgrp(G, i) at f − 2s, f − s, f, f + s; out[i] = its flux */
template <bool LAND, int G>
__device__ __forceinline__ void face_fluxes_grp(const DevGrid& g, const Face& F, bool zero, uint32_t f, uint32_t s,
                                                const float (*v)[4], float* out)
{
    uint32_t need = 0;
#pragma unroll
    for (int i = 0; i < fg_n(G); ++i) {
        const Sten t = stencil_of<LAND, true>(g, f, s, v[i][0], v[i][1], v[i][2], v[i][3]);
        bool bad;
        const float r = flux_fast(F, t, bad);
        out[i] = zero ? 0.0f : r;
        need |= (uint32_t)(bad && !zero) << i;
    }
    if (__builtin_expect(need != 0, 0)) {
#pragma unroll
        for (int i = 0; i < fg_n(G); ++i)
            if ((need >> i) & 1u) out[i] = flux_full(F, stencil_of<LAND, true>(g, f, s, v[i][0], v[i][1], v[i][2], v[i][3]));
    }
}

/* the tracers of group G at one cell: the fluxes of the face below it, the
 * updates, the stores and the carried state (cu, cs, fz, dz; indexed by the
 * position in the group, cs by sink_index) */
template <bool LAND, bool BCS, bool GM, bool FE, int G>
__device__ __forceinline__ void final_grp(const FinalArgs& a, const DevGrid& g, const float* C0, float* OUT,
                                          const Face& Fb, bool zero, const ZCell& z, const float* tt, float* cu,
                                          float* fz, float* dz, float* cs, float (*pw)[4], const HaloImg& hi)
{
    const uint32_t sz = (uint32_t)g.Px * (uint32_t)g.Py, sy = (uint32_t)g.Px;
    const uint32_t cc = z.cc, c = z.c;
    float f[FG_MAX];
    face_fluxes_grp<LAND, G>(g, Fb, zero, c, sz, pw, f);
    float outv[FG_MAX], dv[FG_MAX], cnv[FG_MAX];
#pragma unroll
    for (int i = 0; i < fg_n(G); ++i) {
        const int q = fgrp(G, i);
        const float* co = fld(C0, g, q);
        cnv[i] = ld(co, c);
        const float pk = (has_plk(q) || (FE && (q == PI_DFe || q == PI_PFe_bio))) ? ld(fld(a.plk, g, q), cc) : 0.0f;
        const int si = sink_index(q);
        /* pw[i][1] = the z-sweep input at cc */
        cell_update<LAND, BCS, GM>(a, g, z, q, tt, cu[i], cnv[i], ld(co, cc - 1), ld(co, cc + 1), ld(co, cc - sy),
                          ld(co, cc + sy), pw[i][1], fz[i], f[i], dz[i], si >= 0 ? cs[si] : 0.0f,
                          pk, outv[i], dv[i]);
    }
    /* the halo images of this cell (fill_halo_tracer! of the new tracers): a
     * lane at the domain's edges writes its value to the halo cells that copy it */
    const int sx = hi.sx, sy0 = hi.sy, sz0 = z.k - 1;
    const bool zedge = sz0 < g.Hz + 2 || sz0 >= g.Hz + g.Nz - 2;          /* the same for the whole wave */
    const uint32_t mz = (a.fill_halo && zedge) ? halo_images(sz0, g.Hz, g.Nz, g.tz) : 1u;
    const bool img = (hi.mx | hi.my | mz) != 1u;
#pragma unroll
    for (int i = 0; i < fg_n(G); ++i) {
        const int q = fgrp(G, i);
        st(fld(OUT, g, q), cc, outv[i]);
        fz[i] = f[i]; dz[i] = dv[i];
        const int si = sink_index(q);
        if (si >= 0) cs[si] = cu[i];
        cu[i] = cnv[i];
    }
    if (img) {
        for (uint32_t m3 = mz; m3; m3 &= m3 - 1u) {
            const int c3 = __ffs(m3) - 1, kz = halo_cand(c3, sz0, g.Hz, g.Nz);
            for (uint32_t m2 = hi.my; m2; m2 &= m2 - 1u) {
                const int b3 = __ffs(m2) - 1, ky = halo_cand(b3, sy0, g.Hy, g.Ny);
                for (uint32_t m1 = hi.mx; m1; m1 &= m1 - 1u) {
                    const int a3 = __ffs(m1) - 1, kx = halo_cand(a3, sx, g.Hx, g.Nx);
                    if (!(a3 | b3 | c3)) continue;
                    const uint32_t h = (uint32_t)(kx + g.Px * (ky + g.Py * kz));
#pragma unroll
                    for (int i = 0; i < fg_n(G); ++i) st(fld(OUT, g, fgrp(G, i)), h, outv[i]);
                }
            }
        }
    }
}

/* one column run (cells k0..k1-1 of column i, j) for the tracers of group G */
template <bool LAND, bool BCS, bool GM, bool FE, int G>
__device__ __forceinline__ void final_run(const FinalArgs& a, const DevGrid& g, const float* C0, float* OUT, int i,
                                          int j, int r, int it)
{
    const uint32_t sz = (uint32_t)g.Px * (uint32_t)g.Py;
    const float dt = a.dt;
    const long long m2 = col2(g, i, j);
    const float Az = GM ? g.mAz[m2] : g.dx * g.dy;
    const float kvAz = a.kv * Az;
    const int k0 = g.Hz + 1 + r * a.run;
    const int k1 = min(k0 + a.run, g.Nz + g.Hz + 1);     /* cells k0..k1-1, faces k0..k1 */
    uint32_t c = (uint32_t)gidx(g, i, j, k0);            /* cell below the current face */
    float cu[FG_MAX], cs[3] = {0.0f, 0.0f, 0.0f}, fz[FG_MAX], dz[FG_MAX];
    HaloImg hi;
    hi.mx = a.fill_halo ? halo_images(i - 1, g.Hx, g.Nx, g.tx) : 1u;
    hi.my = a.fill_halo ? halo_images(j - 1, g.Hy, g.Ny, g.ty) : 1u;
    hi.sx = i - 1; hi.sy = j - 1;
    /* z-sweep input around the current face c: pw[i] = prev at c − 2sz, c − sz,
     * c, c + sz (slid up one level per cell, one new load per tracer) */
    float pw[FG_MAX][4];
#pragma unroll
    for (int x = 0; x < fg_n(G); ++x) {
        const float* p = fld(a.prev, g, fgrp(G, x));
        pw[x][0] = ld(p, c - 2 * sz); pw[x][1] = ld(p, c - sz); pw[x][2] = ld(p, c); pw[x][3] = ld(p, c + sz);
    }
    Face Fa = face(g, ld(a.w, c), dt, dzC_at(g, k0), g.ydzC[k0 - 1], Az);
    face_fluxes_grp<LAND, G>(g, Fa, false, c, sz, pw, fz);
    {
        const float dzf = dzF_at(g, k0);
        const double ydzf = g.ydzF[k0 - 1];
#pragma unroll
        for (int x = 0; x < fg_n(G); ++x) {
            const float* co = fld(C0, g, fgrp(G, x));
            const float above = ld(co, c - sz), here = ld(co, c);
            dz[x] = kvAz * divc(lmask<LAND>(g, c - sz, c, above - here), dzf, ydzf);
            cu[x] = here;
        }
        if (G == 3) {
            cs[0] = ld(fld(C0, g, PI_PFe_inorg), c - sz);
            cs[1] = ld(fld(C0, g, PI_Dust), c - sz);
            cs[2] = ld(fld(C0, g, PI_PFe_bio), c - sz);
        }
    }
#pragma unroll 1
    for (int kk = k0 + 1; kk <= k1; ++kk) {
        c += sz;
#pragma unroll
        for (int x = 0; x < fg_n(G); ++x) {
            pw[x][0] = pw[x][1]; pw[x][1] = pw[x][2]; pw[x][2] = pw[x][3];
            pw[x][3] = ld(fld(a.prev, g, fgrp(G, x)), c + sz);
        }
        ZCell z;
        z.c = c; z.cc = c - sz; z.k = kk - 1; z.i = i; z.j = j; z.it = it;
        const Face Fb = face(g, ld(a.w, c), dt, dzC_at(g, kk), g.ydzC[kk - 1], Az);
        const bool zero = kk == g.Nz + g.Hz + 1 && g.tz == PI_BOUNDED;
        z.dT = lmask<LAND>(g, z.cc, c, Fa.T - Fb.T);
        if (GM) {
            const uint32_t cc = z.cc;
            z.vol = g.mVol[cc];
            z.yvol = 0.0;
            z.coef = dt / z.vol;
            z.khAx = a.kh * g.mAx[cc]; z.khAxE = a.kh * g.mAx[cc + 1];
            z.khAy = a.kh * g.mAy[cc]; z.khAyN = a.kh * g.mAy[cc + (uint32_t)g.Px];
            z.dxFW = g.mdxF[m2]; z.dxFE = g.mdxF[m2 + 1];
            z.dyFS = g.mdyF[m2]; z.dyFN = g.mdyF[m2 + g.Px];
            z.kvAz = kvAz;
            z.inv_vol = 1.0f / z.vol;
        } else {
            z.vol = g.vol[z.k - 1];
            z.yvol = g.yvol[z.k - 1];
            z.coef = divc(dt, z.vol, z.yvol);
            z.khAx = a.kh * (g.dy * dzF_at(g, z.k));
            z.khAy = a.kh * (g.dx * dzF_at(g, z.k));
            z.inv_vol = divc(1.0f, z.vol, z.yvol);
        }
        z.dzf = dzF_at(g, kk);
        z.ydzf = g.ydzF[kk - 1];
        float tt[PI_NTRACERS] = {};    /* tracer_temp = max(0, c) (tracer_forcings.jl:3-5) */
#pragma unroll
        for (int x = 0; x < fg_n(G); ++x) tt[fgrp(G, x)] = dmaxf(0.0f, cu[x]);
        if (G == 3) tt[PI_POC] = dmaxf(0.0f, ld(fld(C0, g, PI_POC), z.cc));
        final_grp<LAND, BCS, GM, FE, G>(a, g, C0, OUT, Fb, zero, z, tt, cu, fz, dz, cs, pw, hi);
        Fa = Fb;
    }
}

template <bool LAND, bool BCS, bool GM, bool FE>
__device__ __forceinline__ void k_final_body(const FinalArgs& a, Blk K)
{
    const DevState* S = a.st;
    if (S->err) return;
    const int it = BCS ? (int)(S->iteration + 1) : 0;   /* model.iteration during TimeStep! */
    const DevGrid& g = a.g;
    const int nrun = (g.Nz + a.run - 1) / a.run;
    const long long nthr = (long long)g.Nx * g.Ny * nrun * FG_NGRP;
    const float* C0 = S->tr[S->tcur][0];
    float* OUT = S->tr[S->tcur ^ 1][0];
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < nthr; t += (long long)K.n * blockDim.x) {
        /* t = i + Nx·(group + 4·(j + Ny·run)): the groups of a row are neighbouring blocks */
        const int i = (int)(t % g.Nx) + 1 + g.Hx;
        const long long u = t / g.Nx;
        const int G = (int)(u & 3);
        const int j = (int)((u >> 2) % g.Ny) + 1 + g.Hy;
        const int r = (int)((u >> 2) / g.Ny);
        switch (G) {
            case 0: final_run<LAND, BCS, GM, FE, 0>(a, g, C0, OUT, i, j, r, it); break;
            case 1: final_run<LAND, BCS, GM, FE, 1>(a, g, C0, OUT, i, j, r, it); break;
            case 2: final_run<LAND, BCS, GM, FE, 2>(a, g, C0, OUT, i, j, r, it); break;
            default: final_run<LAND, BCS, GM, FE, 3>(a, g, C0, OUT, i, j, r, it); break;
        }
    }
}
/* at least 5 waves per SIMD: 0.398 -> 0.383 ms on the 256x256x64 run (4 waves at 108 VGPRs) */
/* k_final<n>: bit 0 landmask, bit 1 flux boundary conditions (apply_bcs!),
 * bit 2 LatLonGrid metrics, bit 3 IronEnergyMode (DFe and PFe_bio exchange:
 * a compile-time switch, since the extra loads spill registers of the other
 * variants, 0.39 -> 0.77 ms on the 256x256x64 run when read from a flag) */
#define PI_FINAL(N, LAND, BCS, GM, FE)                                                                \
    __device__ __forceinline__ void k_final##N##_body(const FinalArgs& a, Blk K) { k_final_body<LAND, BCS, GM, FE>(a, K); } \
    PI_KERNEL3W(k_final##N, k_final##N##_body, FinalArgs, 256, xcd_block(), 5)
PI_FINAL(0, false, false, false, false)
PI_FINAL(1, true, false, false, false)
PI_FINAL(2, false, true, false, false)
PI_FINAL(3, true, true, false, false)
PI_FINAL(4, false, false, true, false)
PI_FINAL(5, true, false, true, false)
PI_FINAL(6, false, true, true, false)
PI_FINAL(7, true, true, true, false)
PI_FINAL(8, false, false, false, true)
PI_FINAL(9, true, false, false, true)
PI_FINAL(10, false, true, false, true)
PI_FINAL(11, true, true, false, true)
PI_FINAL(12, false, false, true, true)
PI_FINAL(13, true, false, true, true)
PI_FINAL(14, false, true, true, true)
PI_FINAL(15, true, true, true, true)

/* ------------------------------------------------- fused x + y sweeps (k_xy)
 * calc_Gcsˣ! + multi_dim_x! + fill_halo_tracer! + calc_Gcsʸ! + multi_dim_y!
 * (Advection/multi_dim_adv.jl:111-135) in one pass, then the z halos of the
 * result on interior lines (the halo fill the z sweep reads), for grids whose
 * Periodic directions have N >= H, H = 2 and Nx <= XY_W (blocks of XY_T
 * threads up to Nx = XY_T, of XY_W threads above).
 *
 * A group of roundup64(Nx) threads owns one x-row (thread = cell i) and walks
 * a run of rows (XYArgs::run) of one level in y.  Each row step computes the x sweep
 * of one row: every thread the flux of its west face; the east face comes
 * from the next thread through LDS (one barrier per row, double-buffered); the
 * east face of the last cell is the first face's flux when x is Periodic
 * (the reference's Gc halo copy) and 0 when Bounded.  The x-swept values
 * (temp after the x sweep) of the last rows stay in registers with their
 * masked y differences, so each y face flux and each difference is computed
 * once, and the y update of the row two steps back is written out.
 * Rows outside the interior are the reference's halo fill of the x-swept
 * field: the x sweep of the row they copy (periodic image or Bounded edge),
 * recomputed bit for bit.  Tracers are processed in two passes of 7. */
#define XY_T 256
#define XY_W 512

__device__ __forceinline__ int map_row(int j, int N, int topo)
{
    if (topo == PI_PERIODIC) return j < 0 ? j + N : j >= N ? j - N : j;
    return j < 0 ? 0 : j >= N ? N - 1 : j;
}

/* fast flux from the stencil values bp, bm and the masked differences dd0
 * (at the face), dnp, dnm (the faces before and after, in the x/y sense of
 * stencil()); bad = the full formula is needed */
__device__ __forceinline__ float flux_fast_s(const Face& F, float bp, float bm, float dd0, float dnp, float dnm, bool& bad)
{
    const bool pos = F.T > 0.0f;
    bool sub = false;
    const float p = psi<false>(F, theta(pos ? dnp : dnm, dd0), sub);
    const float pd = p * dd0;
    const float r = F.T * (pos ? bp + pd : bm - pd);
    /* finite stencil values and differences below 2^126: every Ψ is in [0, 1]
     * and the omitted term is finite (NaN fails the comparison) */
    const float m = dmaxf(dmaxf(dmaxf(fabsf(bp), fabsf(bm)), fabsf(dd0)), dmaxf(fabsf(dnp), fabsf(dnm)));
    bad = !(m < 0x1p126f && F.fast && r != 0.0f) || sub;
    return r;
}
__device__ __forceinline__ float flux_full_s(const Face& F, float bp, float bm, float dd0, float dnp, float dnm)
{
    const float T = F.T;
    bool sub = false;
    const float pp = psi<true>(F, theta(dnp, dd0), sub), pm = psi<true>(F, theta(dnm, dd0), sub);
    return 0.5f * (T + fabsf(T)) * (bp + pp * dd0) + 0.5f * (T - fabsf(T)) * (bm - pm * dd0);
}

template <int T>
struct XYShared {
    float c[2][7][T];         /* tracer rows of the x sweep (double-buffered) */
    float F[2][7][T];         /* their west-face fluxes */
};

/* One pass (tracers Q0..Q1-1) over a run of rows of one level.  Row step n
 * (n = 0 .. run+3) is the x sweep of row jr = j0 - 2 + n.  Tracer rows are
 * loaded straight into LDS (global_load_lds: no registers) one row step ahead,
 * each cell once: the x stencil reads its neighbours from LDS.  In the interval
 * between two barriers a thread finishes row n-1 (east face from LDS, the y
 * faces and the update of row n-3), starts the load of row n+1 and computes
 * the fluxes of row n; one barrier per row step.  x halo values are the
 * reference's filled halos of the row: periodic images, or the edge value when
 * Bounded. */
template <bool LAND, int Q0, int Q1, int TB>
__device__ __forceinline__ void xy_run(const XYArgs& a, const DevGrid& g, const PI_GLOBAL float* C0, PI_GLOBAL float* OUT,
                                       XYShared<TB>& sh, int t, int gbase, int i, bool xin, bool act, int j0, int k)
{
    constexpr int NQ = Q1 - Q0;
    const int NS = a.run + 4;
    const uint32_t sy = (uint32_t)g.Px, nc = (uint32_t)g.ncell;
    const PI_GLOBAL float* U = gp(a.u);
    const PI_GLOBAL float* V = gp(a.v);
    const float dt = a.dt;
    const int kk = k + g.Hz + 1;                                  /* 1-based padded level */
    const float Ax = g.dy * dzF_at(g, kk), Ay = g.dx * dzF_at(g, kk);
    const float coef = divc(dt, g.vol[kk - 1], g.yvol[kk - 1]);
    const uint32_t zoff = sy * (uint32_t)g.Py * (uint32_t)(k + g.Hz);
    const uint32_t xoff = (uint32_t)(i + g.Hx);
    auto cellof = [&](int j) -> uint32_t { return xoff + sy * (uint32_t)(j + g.Hy) + zoff; };
    const bool last = i == g.Nx - 1;
    const bool perx = g.tx == PI_PERIODIC;
    /* x stencil neighbours within the row (LDS slots of this group) */
    const int im2 = gbase + (perx ? (i >= 2 ? i - 2 : i - 2 + g.Nx) : max(i - 2, 0));
    const int im1 = gbase + (perx ? (i >= 1 ? i - 1 : i - 1 + g.Nx) : max(i - 1, 0));
    const int ip1 = gbase + (perx ? (i + 1 < g.Nx ? i + 1 : i + 1 - g.Nx) : min(i + 1, g.Nx - 1));
    const int tn = last ? gbase : min(t + 1, TB - 1);            /* slot of the east face's flux */
    float A[NQ], B[NQ], dA[NQ], dB[NQ], Fp[NQ];
    const int wbase = __builtin_amdgcn_readfirstlane(t) & ~63;    /* this wave's LDS slots (uniform) */
    /* row jr of tracers Q0.. into LDS buffer b: each lane loads its own cell */
    auto stage = [&](int b, int jr) {
        const uint32_t c = cellof(map_row(jr, g.Ny, g.ty));
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            __builtin_amdgcn_global_load_lds((PI_GLOBAL void*)((const PI_GLOBAL char*)(C0 + (uint32_t)(Q0 + q) * nc) + c * 4u),
                                             (__attribute__((address_space(3))) void*)&sh.c[b][q][wbase], 4, 0, 0);
    };
    float Tp = 0.0f, Tx = 0.0f, dTx = 0.0f;
    uint32_t cxp = 0;                                             /* cell of the previous row step */
#pragma unroll
    for (int q = 0; q < NQ; ++q) { A[q] = B[q] = dA[q] = dB[q] = Fp[q] = 0.0f; }
    stage(0, j0 - 2);
    __builtin_amdgcn_s_waitcnt(PI_WAIT_VM0);
    __syncthreads();
#pragma unroll 1
    for (int n = 0; n <= NS; ++n) {
        /* ---- finish row step s = n - 1: x update, then y */
        if (n >= 1) {
            const int s = n - 1;
            const int pp = s & 1;
            const int jr = j0 - 2 + s;
            float N[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const float Fo = sh.F[pp][q][t];
                const float Fn = (last && !perx) ? 0.0f : sh.F[pp][q][tn];
                const float co = sh.c[pp][q][t];
                N[q] = co - coef * (lmask<LAND>(g, cxp, cxp + 1, Fn - Fo) - co * dTx);
            }
            /* difference at face jr (between rows jr-1 and jr), with the
             * landmask of the rows themselves (halo rows included) */
            const uint32_t cl = cellof(max(-2, min(jr, g.Ny + 1))), cl1 = cellof(max(-2, min(jr - 1, g.Ny + 1)));
            float dN[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) dN[q] = lmask<LAND>(g, cl1, cl, N[q] - B[q]);
            if (s >= 3) {
                /* face f = jr - 1 (south face of row f): cm2 = row f-2 ... cp1 = row f+1 */
                const int f = jr - 1;
                const uint32_t cf = cellof(min(f, g.Ny));
                const Face Fy = face(g, ldg(V, cf), dt, g.dy, g.ydy, Ay);
                const bool zero = f == g.Ny && g.ty == PI_BOUNDED;
                float G[NQ];
                uint32_t needy = 0;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    bool bad;
                    G[q] = flux_fast_s(Fy, A[q], B[q], dB[q], dA[q], dN[q], bad);
                    needy |= (uint32_t)(bad && !zero) << q;
                }
                if (__builtin_expect(needy != 0, 0)) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
                        if ((needy >> q) & 1u) G[q] = flux_full_s(Fy, A[q], B[q], dB[q], dA[q], dN[q]);
                }
#pragma unroll
                for (int q = 0; q < NQ; ++q) G[q] = zero ? 0.0f : G[q];
                const int jo = f - 1;                             /* row updated by faces f-1 and f */
                if (s >= 4 && act && xin && jo < min(j0 + a.run, g.Ny)) {
                    const uint32_t co = cellof(jo), cn = cellof(f);
                    const float dTy = lmask<LAND>(g, co, cn, Fy.T - Tp);
                    /* z halo images of this cell on its (interior) line: Bounded z copies the edge level */
                    const uint32_t sz = sy * (uint32_t)g.Py;
                    const bool top = k == 0, bot = k == g.Nz - 1;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const PI_GLOBAL float* c = C0 + (uint32_t)(Q0 + q) * nc;
                        PI_GLOBAL float* o = OUT + (uint32_t)(Q0 + q) * nc;
                        const float v = A[q] - coef * (lmask<LAND>(g, co, cn, G[q] - Fp[q]) - ldg(c, co) * dTy);
                        stg(o, co, v);
                        if (top) { stg(o, co - sz, v); stg(o, co - 2 * sz, v); }
                        if (bot) { stg(o, co + sz, v); stg(o, co + 2 * sz, v); }
                    }
                }
#pragma unroll
                for (int q = 0; q < NQ; ++q) Fp[q] = G[q];
                Tp = Fy.T;
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) { A[q] = B[q]; B[q] = N[q]; dA[q] = dB[q]; dB[q] = dN[q]; }
        }
        if (n == NS) break;
        const int p = n & 1;
        /* ---- load row n + 1 into the buffer whose row n - 1 was read above */
        if (n + 1 < NS) stage(p ^ 1, j0 - 1 + n);
        /* ---- x fluxes of row step n, from the staged row */
        const uint32_t cx = cellof(map_row(j0 - 2 + n, g.Ny, g.ty));
        const Face Fx = face(g, ldg(U, cx), dt, g.dx, g.ydx, Ax);
        uint32_t need = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const float cm2 = sh.c[p][q][im2], cm1 = sh.c[p][q][im1], c0 = sh.c[p][q][t], cp1 = sh.c[p][q][ip1];
            const float dd0 = lmask<LAND>(g, cx - 1, cx, c0 - cm1);
            const float dnp = lmask<LAND>(g, cx - 2, cx - 1, cm1 - cm2);
            const float dnm = lmask<LAND>(g, cx, cx + 1, cp1 - c0);
            bool bad;
            const float F = flux_fast_s(Fx, cm1, c0, dd0, dnp, dnm, bad);
            need |= (uint32_t)bad << q;
            sh.F[p][q][t] = F;
        }
        if (__builtin_expect(need != 0, 0)) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if ((need >> q) & 1u) {
                    const float cm2 = sh.c[p][q][im2], cm1 = sh.c[p][q][im1], c0 = sh.c[p][q][t], cp1 = sh.c[p][q][ip1];
                    sh.F[p][q][t] = flux_full_s(Fx, cm1, c0, lmask<LAND>(g, cx - 1, cx, c0 - cm1),
                                                lmask<LAND>(g, cx - 2, cx - 1, cm1 - cm2),
                                                lmask<LAND>(g, cx, cx + 1, cp1 - c0));
                }
            }
        }
        Tx = Fx.T;
        dTx = lmask<LAND>(g, cx, cx + 1, Ax * ldg(U, cx + 1) - Tx);   /* transport difference across the cell */
        cxp = cx;
        __builtin_amdgcn_s_waitcnt(PI_WAIT_VM0);                  /* row n + 1 is in LDS */
        __syncthreads();
    }
}

template <bool LAND, int TB>
__device__ __forceinline__ void k_xy_body(const XYArgs& a, Blk K)
{
    __shared__ XYShared<TB> sh;
    const DevState* S = a.st;
    if (S->err) return;
    const DevGrid& g = a.g;
    const int nxr = (g.Nx + 63) & ~63;
    const int G = TB / nxr;                                       /* x-rows per block */
    const int t = (int)threadIdx.x;
    const int grp = t / nxr;
    const int gbase = grp * nxr;
    const int x = t - gbase;
    const bool xin = grp < G && x < g.Nx;
    const int i = min(x, g.Nx - 1);
    const int nrun = (g.Ny + a.run - 1) / a.run;
    const uint32_t nitem = (uint32_t)nrun * (uint32_t)g.Nz;
    const PI_GLOBAL float* C0 = gp((const float*)S->tr[S->tcur][0]);
    PI_GLOBAL float* OUT = gp(a.out);
    for (uint32_t base = K.b * (uint32_t)G; base < nitem; base += K.n * (uint32_t)G) {
        const uint32_t item = base + (uint32_t)min(grp, G - 1);
        const bool act = grp < G && item < nitem;
        const uint32_t it = act ? item : base;
        const int r = (int)(it % (uint32_t)nrun), k = (int)(it / (uint32_t)nrun);
        xy_run<LAND, 0, 7>(a, g, C0, OUT, sh, t, gbase, i, xin, act, r * a.run, k);
        xy_run<LAND, 7, 14>(a, g, C0, OUT, sh, t, gbase, i, xin, act, r * a.run, k);
    }
}
__device__ __forceinline__ void k_xy0_body(const XYArgs& a, Blk K) { k_xy_body<false, XY_T>(a, K); }
__device__ __forceinline__ void k_xy1_body(const XYArgs& a, Blk K) { k_xy_body<true, XY_T>(a, K); }
__device__ __forceinline__ void k_xyw0_body(const XYArgs& a, Blk K) { k_xy_body<false, XY_W>(a, K); }
__device__ __forceinline__ void k_xyw1_body(const XYArgs& a, Blk K) { k_xy_body<true, XY_W>(a, K); }
/* at least 4 waves per SIMD (128 VGPRs instead of 130-134, which allowed 3):
 * 0.460 -> 0.327 ms on the 256x256x64 run */
PI_KERNEL3W(k_xy0, k_xy0_body, XYArgs, XY_T, xcd_block(), 4)
PI_KERNEL3W(k_xy1, k_xy1_body, XYArgs, XY_T, xcd_block(), 4)
PI_KERNEL3W(k_xyw0, k_xyw0_body, XYArgs, XY_W, xcd_block(), 4)
PI_KERNEL3W(k_xyw1, k_xyw1_body, XYArgs, XY_W, xcd_block(), 4)

bool xy_supported(const DevGrid& g)
{
    return !g.gm && g.Hx == 2 && g.Hy == 2 && g.Hz == 2 && g.Nx <= XY_W && g.tz == PI_BOUNDED &&
           !(g.tx == PI_PERIODIC && g.Nx < g.Hx) && !(g.ty == PI_PERIODIC && g.Ny < g.Hy);
}
bool xy_wide(const DevGrid& g) { return g.Nx > XY_T; }
long long xy_blocks(const DevGrid& g, int run)
{
    const int nxr = (g.Nx + 63) & ~63;
    const long long G = (xy_wide(g) ? XY_W : XY_T) / nxr;
    const long long nitem = (long long)((g.Ny + run - 1) / run) * g.Nz;
    return (nitem + G - 1) / G;
}

/* ------------------------------------------------------------- registry */
long long sweep_x_blocks(const DevGrid& g)
{
    const long long nface = (long long)g.Ny * g.Nz * (g.Nx + 1);
    const long long nwave = (nface + 62) / 63;      /* 63 faces' updates per wave */
    return (nwave + 3) / 4;
}
long long sweep_y_items(const DevGrid& g) { return (long long)g.Nx * ((g.Ny + RUN_Y - 1) / RUN_Y) * g.Nz; }
long long final_items(const DevGrid& g, int run) { return (long long)g.Nx * g.Ny * ((g.Nz + run - 1) / run) * FG_NGRP; }

void reg_tracers(KInfo* k)
{
    PI_REG(k, K_SWEEPX0, k_sweep_x0, SweepArgs, 256, false);
    PI_REG(k, K_SWEEPX1, k_sweep_x1, SweepArgs, 256, false);
    PI_REG(k, K_SWEEPX2, k_sweep_x2, SweepArgs, 256, false);
    PI_REG(k, K_SWEEPX3, k_sweep_x3, SweepArgs, 256, false);
    PI_REG(k, K_SWEEPY0, k_sweep_y0, SweepArgs, 256, true);
    PI_REG(k, K_SWEEPY1, k_sweep_y1, SweepArgs, 256, true);
    PI_REG(k, K_SWEEPY2, k_sweep_y2, SweepArgs, 256, true);
    PI_REG(k, K_SWEEPY3, k_sweep_y3, SweepArgs, 256, true);
    PI_REG(k, K_FILL, k_fill_line, FillArgs, 256, true);
    PI_REG(k, K_HALOCOPY, k_halo_copy, FillArgs, 256, true);
    PI_REG(k, K_FINAL0, k_final0, FinalArgs, 256, true);
    PI_REG(k, K_FINAL1, k_final1, FinalArgs, 256, true);
    PI_REG(k, K_FINAL2, k_final2, FinalArgs, 256, true);
    PI_REG(k, K_FINAL3, k_final3, FinalArgs, 256, true);
    PI_REG(k, K_FINAL4, k_final4, FinalArgs, 256, true);
    PI_REG(k, K_FINAL5, k_final5, FinalArgs, 256, true);
    PI_REG(k, K_FINAL6, k_final6, FinalArgs, 256, true);
    PI_REG(k, K_FINAL7, k_final7, FinalArgs, 256, true);
    PI_REG(k, K_FINAL8, k_final8, FinalArgs, 256, true);
    PI_REG(k, K_FINAL9, k_final9, FinalArgs, 256, true);
    PI_REG(k, K_FINAL10, k_final10, FinalArgs, 256, true);
    PI_REG(k, K_FINAL11, k_final11, FinalArgs, 256, true);
    PI_REG(k, K_FINAL12, k_final12, FinalArgs, 256, true);
    PI_REG(k, K_FINAL13, k_final13, FinalArgs, 256, true);
    PI_REG(k, K_FINAL14, k_final14, FinalArgs, 256, true);
    PI_REG(k, K_FINAL15, k_final15, FinalArgs, 256, true);
    PI_REG(k, K_XY0, k_xy0, XYArgs, XY_T, true);
    PI_REG(k, K_XY1, k_xy1, XYArgs, XY_T, true);
    PI_REG(k, K_XYW0, k_xyw0, XYArgs, XY_W, true);
    PI_REG(k, K_XYW1, k_xyw1, XYArgs, XY_W, true);
}
