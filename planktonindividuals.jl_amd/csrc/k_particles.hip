/*
 * k_particles.hip - particle kernels of the hot path on gfx950.
 *
 *   k_adv       RK4 advection + random-walk diffusion + cell index, one lane per
 *               live individual, store order (cell-sorted from the previous
 *               step).  An individual that keeps its cell is a STAYER; the
 *               others (movers, and daughters appended last step) are queued.
 *   k_cellcount per cell: stayers (from the scan of stayer flags over the
 *               previous segment) + queued arrivals.
 *   k_place_stay  stayers keep their relative order: scattered without atomics.
 *   k_place_small queued individuals go after the stayers of their cell; such
 *               cells are listed and re-sorted (k_mixsort, k_bigcells).
 *   k_cellsum   per cell: pop = count, Chl = ordered sum (the reference's
 *               atomic order under julia -t 1).
 *   k_*_long    cells with more than LONG_SEG entries (k_cellcount lists
 *               them): one block per cell, the entries staged through LDS in
 *               chunks and the ordered sums run by one wave, so the chain of
 *               additions runs at the add latency instead of a load each.
 *   k_marks     survivors/dividers as per-species slot bitmaps, from per-slot
 *               flag bytes written by k_phys (no atomics).
 *   k_par       Beer-Lambert column scan (PAR), one lane per column.
 *   k_phys      per individual in cell order: losses, division (slot assignment
 *               from the bitmaps), QuotaMode growth, decisions, consumption.
 *   k_plk       per cell: ordered sums of losses/consumption into plk + counts
 *               for the diagnostics.
 *
 * Every floating-point expression keeps the reference's operation order
 * (cited per function); the build uses -ffp-contract=off and no fast-math.
 */
#include "pi_device.h"
#include "k_args.h"

#define SMALL_SEG 16
#define LONG_SEG 64       /* longer cells take the one-block-per-cell kernels */
#define LCH 1024          /* entries per staged chunk of those kernels */
#define LPAD (LCH + 4)    /* row stride of the staged values (rows start in different LDS banks) */
#define MAXARR 2048       /* arrivals of a re-sorted cell sorted in LDS */

/* ---------------------------------------------------------------- advection
 * u_itpl / v_itpl / w_itpl with ac = true (Individuals/Advection/interpolation.jl:1-83).
 * Corner indices are clamped into the padded array so that positions outside
 * the halo (not produced by the reference's boundary conditions) cannot fault.
 * The corners (x, x+1) are read in pairs, four 8-byte loads per call.  Measured
 * on MI355X and not kept: a 5-instruction exact quotient (13% fewer VALU
 * instructions: slower), re-dealing lanes by corner rows (fewer cache lines per
 * gather: no change), staging each block's velocity neighbourhood in LDS (the
 * texture path relieved, 20% slower). */
template <bool GM>
__device__ __forceinline__ float itpl(const DevGrid& g, const float* __restrict__ f,
                                      float x, float y, float z, int comp)
{
    float xs = comp == 0 ? x + 1.0f : x + 0.5f;
    float ys = comp == 1 ? y + 1.0f : y + 0.5f;
    float zs = comp == 2 ? z + 1.0f : z + 0.5f;
    float xd = jl_mod1f(xs), yd = jl_mod1f(ys), zd = jl_mod1f(zs);
    int xi = (int)xs + g.Hx, yi = (int)ys + g.Hy, zi = (int)zs + g.Hz;
    xi = min(max(xi, 1), g.Px - 1);
    yi = min(max(yi, 1), g.Py - 1);
    zi = min(max(zi, 1), g.Pz - 1);
    const uint32_t sx = 1, sy = (uint32_t)g.Px, sz = (uint32_t)g.Px * (uint32_t)g.Py;
    const uint32_t b = (uint32_t)(xi - 1) + sy * (uint32_t)(yi - 1) + sz * (uint32_t)(zi - 1);
    float D0, D1;             /* divisors and their reciprocals */
    double R0, R1;
    if (comp == 0) { D0 = g.dx; D1 = g.dx; R0 = g.ydx; R1 = g.ydx; }
    else if (comp == 1) { D0 = g.dy; D1 = g.dy; R0 = g.ydy; R1 = g.ydy; }
    else { D0 = dzC_at(g, zi); D1 = dzC_at(g, zi + 1); R0 = g.ydzC[zi - 1]; R1 = g.ydzC[zi]; }
    const PI_GLOBAL float* F = gp(f);
    const pi_f2u p00 = ldg2(F, b), p01 = ldg2(F, b + sz), p10 = ldg2(F, b + sy), p11 = ldg2(F, b + sy + sz);
    const float f000 = p00.x, f001 = p01.x, f010 = p10.x, f011 = p11.x;
    const float f100 = p00.y, f101 = p01.y, f110 = p10.y, f111 = p11.y;
    (void)sx;
    const float ox = 1.0f - xd, oy = 1.0f - yd, oz = 1.0f - zd;
    if (GM && comp < 2) {
        /* LatLonGrid: ΔxC / ΔyC of each corner column (the z corners share it) */
        const float* M = comp == 0 ? g.mdxC : g.mdyC;
        const long long m = col2(g, xi, yi);
        const float E00 = M[m], E01 = M[m + g.Px], E10 = M[m + 1], E11 = M[m + 1 + g.Px];
        float s = (ox * oy * oz * f000) / E00;
        s = s + (ox * oy * zd * f001) / E00;
        s = s + (ox * yd * oz * f010) / E01;
        s = s + (ox * yd * zd * f011) / E01;
        s = s + (xd * oy * oz * f100) / E10;
        s = s + (xd * oy * zd * f101) / E10;
        s = s + (xd * yd * oz * f110) / E11;
        s = s + (xd * yd * zd * f111) / E11;
        return s;
    }
    float s = divc(ox * oy * oz * f000, D0, R0);
    s = s + divc(ox * oy * zd * f001, D1, R1);
    s = s + divc(ox * yd * oz * f010, D0, R0);
    s = s + divc(ox * yd * zd * f011, D1, R1);
    s = s + divc(xd * oy * oz * f100, D0, R0);
    s = s + divc(xd * oy * zd * f101, D1, R1);
    s = s + divc(xd * yd * oz * f110, D0, R0);
    s = s + divc(xd * yd * zd * f111, D1, R1);
    return s;
}

/* particle_boundary_condition (particle_advection_kernels.jl:2-11), as
 * selects: both arms are a subtraction and an addition */
__device__ __forceinline__ float pbc(float x, int N, int topo)
{
    const float xl = 0.0f, xr = (float)N;
    float hi, lo;
    if (topo == PI_PERIODIC) {
        hi = xl + (x - xr);
        lo = xr - (xl - x);
    } else {
        hi = xr - (x - xr) * 1.0f;
        lo = xl + (xl - x) * 1.0f;
    }
    return x > xr ? hi : x < xl ? lo : x;
}


/* one slot of a device queue per lane with pred, one atomic per wave */
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred)
{
    const unsigned long long m = __ballot(pred);
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader < 0 ? 0 : leader, 64);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

/* RK4 advection (particle_advection.jl:2-56) and the random walk
 * (particle_diffusion.jl:2-33) of one individual, up to its new position. */
template <bool GM>
__device__ __forceinline__ void move(const AdvArgs& a, const DevGrid& g, float x0, float y0, float z0,
                                     float r0, float r1, float r2, long long xprev, long long yprev, long long zprev,
                                     float& x, float& y, float& z)
{
    const float dt = a.dt;
    /* ---- RK4, particle_advection.jl:2-56 */
    float u1 = itpl<GM>(g, a.v0.u, x0, y0, z0, 0);
    float v1 = itpl<GM>(g, a.v0.v, x0, y0, z0, 1);
    float w1 = itpl<GM>(g, a.v0.w, x0, y0, z0, 2);
    float u2 = u1, v2 = v1, w2 = w1;
    float px = pbc(x0 + 0.5f * u1 * dt, g.Nx, g.tx);
    float py = pbc(y0 + 0.5f * v1 * dt, g.Ny, g.ty);
    float pz = pbc(z0 - 0.5f * w1 * dt, g.Nz, g.tz);
    u1 = itpl<GM>(g, a.vh.u, px, py, pz, 0);
    v1 = itpl<GM>(g, a.vh.v, px, py, pz, 1);
    w1 = itpl<GM>(g, a.vh.w, px, py, pz, 2);
    u2 = u2 + u1 * 2.0f; v2 = v2 + v1 * 2.0f; w2 = w2 + w1 * 2.0f;
    px = pbc(x0 + 0.5f * u1 * dt, g.Nx, g.tx);
    py = pbc(y0 + 0.5f * v1 * dt, g.Ny, g.ty);
    pz = pbc(z0 - 0.5f * w1 * dt, g.Nz, g.tz);
    u1 = itpl<GM>(g, a.vh.u, px, py, pz, 0);
    v1 = itpl<GM>(g, a.vh.v, px, py, pz, 1);
    w1 = itpl<GM>(g, a.vh.w, px, py, pz, 2);
    u2 = u2 + u1 * 2.0f; v2 = v2 + v1 * 2.0f; w2 = w2 + w1 * 2.0f;
    px = pbc(x0 + 1.0f * u1 * dt, g.Nx, g.tx);
    py = pbc(y0 + 1.0f * v1 * dt, g.Ny, g.ty);
    pz = pbc(z0 - 1.0f * w1 * dt, g.Nz, g.tz);
    u1 = itpl<GM>(g, a.v1.u, px, py, pz, 0);
    v1 = itpl<GM>(g, a.v1.v, px, py, pz, 1);
    w1 = itpl<GM>(g, a.v1.w, px, py, pz, 2);
    u2 = u2 + u1; v2 = v2 + v1; w2 = w2 + w1;
    u2 = divc(u2, 6.0f, g.y6); v2 = divc(v2, 6.0f, g.y6); w2 = divc(w2, 6.0f, g.y6);
    x = pbc(x0 + 1.0f * u2 * dt, g.Nx, g.tx);
    y = pbc(y0 + 1.0f * v2 * dt, g.Ny, g.ty);
    z = pbc(z0 - 1.0f * w2 * dt, g.Nz, g.tz);
    /* ---- random walk, particle_diffusion.jl:2-33 (stale zi from the previous find_inds) */
    r0 = (r0 - 0.5f) * 2.0f; r1 = (r1 - 0.5f) * 2.0f; r2 = (r2 - 0.5f) * 2.0f;
    if (GM) {           /* ΔxC, ΔyC at the stale (xi + Hx, yi + Hy), kept inside the array */
        long long kx = xprev + g.Hx, ky = yprev + g.Hy;
        kx = kx < 1 ? 1 : kx > g.Px ? g.Px : kx;
        ky = ky < 1 ? 1 : ky > g.Py ? g.Py : ky;
        const long long m = col2(g, (int)kx, (int)ky);
        r0 = (r0 * a.sx) / g.mdxC[m];
        r1 = (r1 * a.sy) / g.mdyC[m];
    } else {
        r0 = divc(r0 * a.sx, g.dx, g.ydx);
        r1 = divc(r1 * a.sy, g.dy, g.ydy);
    }
    {
        long long kz = zprev + g.Hz;
        kz = kz < 1 ? 1 : kz > g.Pz ? g.Pz : kz;
        r2 = divc(r2 * a.sz, g.dzC[kz - 1], g.ydzC[kz - 1]);
    }
    x = pbc(x + r0, g.Nx, g.tx);
    y = pbc(y + r1, g.Ny, g.ty);
    z = pbc(z + r2, g.Nz, g.tz);
}

/* One store entry after its move: cell, stayer flag or queue slot, death
 * bookkeeping.  Every lane of the wave calls it (wave_append); `live` = the
 * entry is active and was moved. */
__device__ __forceinline__ void place_entry(const AdvArgs& a, DevState* st, const GStore& S, uint32_t e, uint32_t nA,
                                            int32_t iteration, bool live, uint8_t fl, float x, float y, float z)
{
    const DevGrid& g = a.g;
    bool queue = false;
    uint32_t c = 0;
    if (live) {
        const uint32_t key = S.key[e];
        const int sp = (int)(key >> PI_SLOT_BITS);
        const uint32_t slot = key & PI_SLOT_MASK;
        /* ---- find_inds (Individuals/utils.jl:2-12) */
        int xi = (int)(x + 1.0f) + g.Hx, yi = (int)(y + 1.0f) + g.Hy, zi = (int)(z + 1.0f) + g.Hz;
        xi = min(max(xi, 1), g.Px); yi = min(max(yi, 1), g.Py); zi = min(max(zi, 1), g.Pz);
        c = (uint32_t)gidx(g, xi, yi, zi);
        const uint32_t cprev = S.cell[e];
        S.x[e] = x; S.y[e] = y; S.z[e] = z;
        S.flags[e] = fl & (uint8_t)~F_ZERO;
        a.newc[e] = c;                      /* S.cell keeps the cell of the previous step */
        const bool stay = e < nA && c == cprev;
        if (e < nA) a.stayf[e] = stay ? 1u : 0u;
        queue = !stay;
        if (fl & (F_GRAZ | F_MORT)) {       /* dies this step: remember where, by slot */
            const long long r = (long long)sp * a.cap + slot;
            a.dstamp[r] = iteration;
            a.dcell[r] = c;
            if (a.dbd) a.dbd[r] = a.dbd_ptc ? S.ex[7][e] : S.Cq[e];   /* CarbonMode: the Bd, IronEnergyMode:
                                                                        * the ptc a daughter born there inherits */
        }
    }
    const uint32_t sq = (e >> 8) & (PI_NQ - 1);       /* the same for every lane of a wave */
    const uint32_t qi = wave_append(&st->nsmall[sq], queue);
    if (queue) {
        a.small[sq * a.qcap + qi] = e;
        atomicAdd(&a.cnts[c], 1u);
    }
}

/* the move of entry e from its stored state in step `iteration`.  (Values of
 * the device state are read before the loops that call these: a load that may
 * alias the loop's stores would be a vector load waiting for every outstanding
 * store.) */
template <bool GM>
__device__ __forceinline__ void move_entry(const AdvArgs& a, const GStore& S, uint32_t e, long long iteration, uint8_t fl,
                                           float& x, float& y, float& z)
{
    const DevGrid& g = a.g;
    const uint32_t key = S.key[e];
    const int sp = (int)(key >> PI_SLOT_BITS);
    const uint32_t slot = key & PI_SLOT_MASK;
    float r0, r1, r2;
    pi_uniform3(a.seed, iteration, sp, PI_KIND_DIFFUSION, slot, &r0, &r1, &r2);
    const uint32_t cprev = S.cell[e];
    /* the previous step's find_inds (1-based padded), (0, 0, 0) before the first */
    const long long zprev = (fl & F_ZERO) ? 0 : (long long)(cprev / (uint32_t)(g.Px * g.Py)) + 1;
    const long long xprev = (fl & F_ZERO) ? 0 : (long long)(cprev % (uint32_t)g.Px) + 1;
    const long long yprev = (fl & F_ZERO) ? 0 : (long long)((cprev / (uint32_t)g.Px) % (uint32_t)g.Py) + 1;
    move<GM>(a, g, S.x[e], S.y[e], S.z[e], r0, r1, r2, xprev, yprev, zprev, x, y, z);
}

template <bool GM>
__device__ __forceinline__ void k_adv_impl(const AdvArgs& a, Blk K)
{
    DevState* st = a.st;
    if (st->err) return;
    const GStore S = gstore(st->store[st->scur]);
    const uint32_t n = st->nstore;
    const uint32_t nA = st->nA;
    const long long iteration = st->iteration + 1;
    const uint32_t nloop = (n + 63u) & ~63u;          /* whole waves stay in the loop (wave_append) */
    for (uint32_t e = K.b * blockDim.x + threadIdx.x; e < nloop; e += K.n * blockDim.x) {
        const uint8_t fl = e < n ? S.flags[e] : (uint8_t)0;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (fl & F_AC) move_entry<GM>(a, S, e, iteration, fl, x, y, z);
        place_entry(a, st, S, e, nA, (int32_t)iteration, (fl & F_AC) != 0, fl, x, y, z);
        if (!(fl & F_AC) && e < nA) a.stayf[e] = 0u;
    }
}
/* at least 5 waves per SIMD (88 VGPRs, no spill): the gathers of the corner
 * values wait less; 0.467 -> 0.436 ms on the 256x256x64 run (4 waves at 105) */
__device__ __forceinline__ void k_adv_body(const AdvArgs& a, Blk K) { k_adv_impl<false>(a, K); }
__device__ __forceinline__ void k_adv_gm_body(const AdvArgs& a, Blk K) { k_adv_impl<true>(a, K); }
PI_KERNEL3W(k_adv, k_adv_body, AdvArgs, 256, xcd_block(), 5)
/* LatLonGrid: the spacings of each corner column from the metric arrays */
PI_KERNEL3W(k_adv_gm, k_adv_gm_body, AdvArgs, 256, xcd_block(), 5)

/* Survivors of this step's losses and the dividers, as per-species slot bitmaps
 * (plankton_update.jl:5-24: ac after grazing!/mortality!, dvid .*= ac).  They
 * depend only on the decisions taken at the end of the previous step, kept as
 * one flag byte per slot, so the capacity check runs before anything is
 * modified.  One lane per 32-slot word. */
__device__ __forceinline__ void k_marks_impl(Blk K, const DevState* st, const uint8_t* sflag, uint32_t* bmA, uint32_t* bmD,
                                               int nsp, int nw, long long cap)
{
    if (st->err) return;
    const long long nwords = (long long)nsp * nw;
    for (long long w = K.b * (long long)blockDim.x + threadIdx.x; w < nwords; w += (long long)K.n * blockDim.x) {
        const long long sp = w / nw, wd = w - sp * nw;
        const long long s0 = wd * 32, s1 = s0 + 32 < cap ? s0 + 32 : cap;
        const uint8_t* f = sflag + sp * cap;
        uint32_t A = 0, D = 0;
        if (s1 - s0 == 32 && ((uintptr_t)(f + s0) & 15u) == 0) {
            /* the word's 32 flag bytes as two 16-byte loads */
            const uint4 lo = *(const uint4*)(f + s0), hi = *(const uint4*)(f + s0 + 16);
            const uint32_t q[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int k = 0; k < 32; ++k) {
                const uint32_t b = (q[k >> 2] >> ((k & 3) * 8)) & 0xFFu;
                const bool alive = (b & F_AC) && !(b & (F_GRAZ | F_MORT));
                A |= (uint32_t)alive << k;
                D |= (uint32_t)(alive && (b & F_DVID)) << k;
            }
        } else {
            for (long long sl = s0; sl < s1; ++sl) {
                const uint8_t b = f[sl];
                const bool alive = (b & F_AC) && !(b & (F_GRAZ | F_MORT));
                A |= (uint32_t)alive << (sl - s0);
                D |= (uint32_t)(alive && (b & F_DVID)) << (sl - s0);
            }
        }
        bmA[w] = A;
        bmD[w] = D;
    }
}
__device__ __forceinline__ void k_marks_body(const MarksArgs& x, Blk K) { k_marks_impl(K, x.st, x.sflag, x.bmA, x.bmD, x.nsp, x.nw, x.cap); }
PI_KERNEL(k_marks, MarksArgs, 256)

/* per cell: stayers (their count from the stayer prefix P over the cell's
 * previous segment segp) + queued arrivals */
__device__ __forceinline__ void k_cellcount_impl(Blk K, DevState* st, long long ncell, const uint32_t* segp,
                                                   const uint32_t* P, const uint32_t* cnts, uint32_t* cnt,
                                                   uint32_t* lng)
{
    if (st->err) return;
    const bool have = st->nA > 0;
    const long long nloop = (ncell + 63) & ~63ll;     /* whole waves: wave_append */
    for (long long c = K.b * (long long)blockDim.x + threadIdx.x; c < nloop; c += (long long)K.n * blockDim.x) {
        bool isl = false;
        if (c < ncell) {
            const uint32_t n = (have ? P[segp[c + 1]] - P[segp[c]] : 0u) + cnts[c];
            cnt[c] = n;
            isl = n > LONG_SEG;
        }
        const uint32_t li = wave_append(&st->nlong, isl);
        if (isl) lng[li] = (uint32_t)c;
    }
}
__device__ __forceinline__ void k_cellcount_body(const BinArgs& x, Blk K) { k_cellcount_impl(K, x.st, x.ncell, x.segp, x.P, x.cnts, x.cnt, x.lng); }
PI_KERNEL_XCD(k_cellcount, BinArgs, 256)

/* stayers: the stayers of a cell keep their (species, slot) order; position =
 * start of the cell + rank among the cell's stayers */
__device__ __forceinline__ void k_place_stay_impl(Blk K, const DevState* st, const uint32_t* seg, const uint32_t* segp,
                                                    const uint32_t* P, const uint32_t* stayf, const uint32_t* newc,
                                                    uint32_t* pkey, uint32_t* psrc)
{
    if (st->err) return;
    const GStore S = gstore(st->store[st->scur]);
    const uint32_t nA = st->nA;
    for (uint32_t e = K.b * blockDim.x + threadIdx.x; e < nA; e += K.n * blockDim.x) {
        if (!stayf[e]) continue;
        const uint32_t c = newc[e];
        const uint32_t pos = seg[c] + (P[e] - P[segp[c]]);
        pkey[pos] = S.key[e];
        psrc[pos] = e;
    }
}
__device__ __forceinline__ void k_place_stay_body(const BinArgs& x, Blk K) { k_place_stay_impl(K, x.st, x.seg, x.segp, x.P, x.stayf, x.newc, x.pkey, x.psrc); }
PI_KERNEL_XCD(k_place_stay, BinArgs, 256)

/* exclusive prefix over the 256 threads of a block of v; *tot = the block total */
__device__ __forceinline__ uint32_t block_excl(uint32_t v, uint32_t* ws, uint32_t& tot)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t i = 0; i < w; ++i) off += ws[i];
    tot = ws[0] + ws[1] + ws[2] + ws[3];
    __syncthreads();
    return off + x - v;
}

/* exclusive prefix of the PI_NQ shard counts into pre[0..PI_NQ] (pre[PI_NQ] =
 * total), by every thread of the block */
__device__ __forceinline__ void shard_prefix(const uint32_t* cnt, uint32_t* pre, uint32_t* ws)
{
    static_assert(PI_NQ <= 256, "one shard count per thread of the 256-thread block");
    uint32_t tot;
    const uint32_t p = block_excl(threadIdx.x < PI_NQ ? cnt[threadIdx.x] : 0u, ws, tot);
    if (threadIdx.x < PI_NQ) pre[threadIdx.x] = p;
    if (threadIdx.x == 0) pre[PI_NQ] = tot;
    __syncthreads();
}
/* the shard q of list position t: pre[q] <= t < pre[q + 1] */
__device__ __forceinline__ uint32_t shard_of(const uint32_t* pre, uint32_t t)
{
    uint32_t lo = 0, hi = PI_NQ;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}
/* queued individuals: after the stayers of their cell, in any order; the
 * cell is listed for re-sorting once (by the last one placed) */
__device__ __forceinline__ void k_place_small_impl(Blk K, DevState* st, const uint32_t* seg, const uint32_t* small,
                                                     uint32_t qcap, const uint32_t* newc, uint32_t* cnts, uint32_t* pkey,
                                                     uint32_t* psrc, uint32_t* mix)
{
    if (st->err) return;
    const GStore S = gstore(st->store[st->scur]);
    /* queue position t -> shard q (s_pre[q] <= t < s_pre[q + 1]) and its entry */
    __shared__ uint32_t s_pre[PI_NQ + 1], s_ws[4];
    shard_prefix(st->nsmall, s_pre, s_ws);
    const uint32_t ns = s_pre[PI_NQ];
    const uint32_t nloop = (ns + 63u) & ~63u;
    for (uint32_t t = K.b * blockDim.x + threadIdx.x; t < nloop; t += K.n * blockDim.x) {
        bool last = false;
        uint32_t c = 0;
        if (t < ns) {
            const uint32_t lo = shard_of(s_pre, t);
            const uint32_t e = small[lo * qcap + (t - s_pre[lo])];
            c = newc[e];
            const uint32_t r = atomicSub(&cnts[c], 1u);       /* cnts returns to zero */
            const uint32_t pos = seg[c + 1] - r;
            pkey[pos] = S.key[e];
            psrc[pos] = e;
            last = r == 1u;
        }
        const uint32_t sq = (t >> 8) & (PI_NQ - 1);    /* the same for every lane of a wave */
        const uint32_t mi = wave_append(&st->nmix[sq], last);
        if (last) mix[sq * qcap + mi] = c;
    }
}
__device__ __forceinline__ void k_place_small_body(const BinArgs& x, Blk K) { k_place_small_impl(K, x.st, x.seg, x.small, x.qcap, x.newc, x.cnts, x.pkey, x.psrc, x.mix); }
PI_KERNEL(k_place_small, BinArgs, 256)

/* re-sort the listed cells by key (species-major, then slot): insertion sort
 * for short segments, the rest go to k_bigcells */
__device__ __forceinline__ void k_mixsort_impl(Blk K, DevState* st, const uint32_t* seg, const uint32_t* mix,
                                                 uint32_t qcap, uint32_t* pkey, uint32_t* psrc, uint32_t* big)
{
    if (st->err) return;
    __shared__ uint32_t s_pre[PI_NQ + 1], s_ws[4];
    shard_prefix(st->nmix, s_pre, s_ws);
    const uint32_t nm = s_pre[PI_NQ];
    for (uint32_t m = K.b * blockDim.x + threadIdx.x; m < nm; m += K.n * blockDim.x) {
        const uint32_t q = shard_of(s_pre, m);
        const uint32_t c = mix[q * qcap + (m - s_pre[q])];
        const uint32_t s0 = seg[c], n = seg[c + 1] - s0;
        if (n > SMALL_SEG) { big[atomicAdd(&st->nbig, 1u)] = c; continue; }
        for (uint32_t a = 1; a < n; ++a) {
            uint32_t k = pkey[s0 + a], v = psrc[s0 + a];
            uint32_t b = a;
            while (b > 0 && pkey[s0 + b - 1] > k) {
                pkey[s0 + b] = pkey[s0 + b - 1];
                psrc[s0 + b] = psrc[s0 + b - 1];
                --b;
            }
            pkey[s0 + b] = k;
            psrc[s0 + b] = v;
        }
    }
}
__device__ __forceinline__ void k_mixsort_body(const BinArgs& x, Blk K) { k_mixsort_impl(K, x.st, x.seg, x.mix, x.qcap, x.pkey, x.psrc, x.big); }
PI_KERNEL(k_mixsort, BinArgs, 256)

/* the first position in the key-sorted range [lo, hi) whose key is >= k */
__device__ __forceinline__ uint32_t key_lower_bound(const uint32_t* pkey, uint32_t lo, uint32_t hi, uint32_t k)
{
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pkey[mid] < k) lo = mid + 1; else hi = mid;
    }
    return lo;
}
/* individuals of species sp in a cell's segment [s0, s1) (sorted by key,
 * species-major): what acc_counts! leaves in pop for that species alone
 * when shared_graz != 1 (time_step.jl:98-102: pop zeroed per species) */
__device__ __forceinline__ uint32_t species_count(const uint32_t* pkey, uint32_t s0, uint32_t s1, uint32_t sp)
{
    const uint32_t a = key_lower_bound(pkey, s0, s1, sp << PI_SLOT_BITS);
    const uint32_t b = key_lower_bound(pkey, a, s1, (sp + 1u) << PI_SLOT_BITS);
    return b - a;
}

/* acc_chl! / acc_counts! (Individuals/utils.jl:33-52) as an ordered segmented
 * sum.  pop_sp >= 0 (shared_graz != 1): pop holds the counts of that species
 * only, as the reference's pop after its last species (time_step.jl:98-102). */
__device__ __forceinline__ void k_cellsum_impl(Blk K, const DevState* st, long long ncell, const uint32_t* seg,
                                                 const uint32_t* psrc, const uint32_t* pkey, int pop_sp,
                                                 float* Chl, float* pop)
{
    if (st->err) return;
    const PI_GLOBAL float* chl_store = gp(st->store[st->scur].Chl);
    for (long long c = K.b * (long long)blockDim.x + threadIdx.x; c < ncell; c += (long long)K.n * blockDim.x) {
        const uint32_t s0 = seg[c], n = seg[c + 1] - s0;
        if (n > LONG_SEG) continue;                  /* k_cellsum_long */
        float sum = 0.0f;
        for (uint32_t a = 0; a < n; ++a) sum = sum + chl_store[psrc[s0 + a]];
        pop[c] = (float)(pop_sp < 0 ? n : species_count(pkey, s0, s0 + n, (uint32_t)pop_sp));
        Chl[c] = sum;
    }
}
__device__ __forceinline__ void k_cellsum_body(const BinArgs& x, Blk K) { k_cellsum_impl(K, x.st, x.ncell, x.seg, x.psrc, x.pkey, x.pop_sp, x.Chl, x.pop); }
PI_KERNEL_XCD(k_cellsum, BinArgs, 256)

/* ordered sum of m staged values (row, 16-byte aligned) onto s: blocks of 32
 * values, the LDS reads of the next block issued before the 32 additions of
 * the current one */
__device__ __forceinline__ float add32(float s, const float4 (&v)[8])
{
#pragma unroll
    for (int q = 0; q < 8; ++q) { s = s + v[q].x; s = s + v[q].y; s = s + v[q].z; s = s + v[q].w; }
    return s;
}
__device__ __forceinline__ float chain_sum(float s, const float* row, uint32_t m)
{
    const float4* r4 = (const float4*)row;
    const uint32_t nb = m >> 5;
    if (nb) {
        float4 cur[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) cur[q] = r4[q];
        for (uint32_t b = 1; b < nb; ++b) {
            float4 nxt[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) nxt[q] = r4[b * 8 + q];
            s = add32(s, cur);
#pragma unroll
            for (int q = 0; q < 8; ++q) cur[q] = nxt[q];
        }
        s = add32(s, cur);
    }
    for (uint32_t i = nb << 5; i < m; ++i) s = s + row[i];
    return s;
}

/* gathers the Chl values of entries [from, min(from + LCH, end)) into dst, U
 * per thread (thread t0 + u·step), every load of a thread issued at once */
template <int U>
__device__ __forceinline__ void stage_chl(const PI_GLOBAL float* chl, const uint32_t* psrc, uint32_t from, uint32_t end,
                                          uint32_t t0, uint32_t step, float* dst)
{
    uint32_t src[U];
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t i = t0 + u * step;
        src[u] = i < LCH && from + i < end ? psrc[from + i] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = src[u] != 0xFFFFFFFFu ? chl[src[u]] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (src[u] != 0xFFFFFFFFu) dst[t0 + u * step] = v[u];
}

/* k_cellsum for the listed long cells: the Chl values of a chunk are gathered
 * into LDS by waves 1-3 while wave 0 adds up the previous chunk */
__device__ __forceinline__ void k_cellsum_long_impl(Blk K, const DevState* st, const uint32_t* lng, const uint32_t* seg,
                                                      const uint32_t* psrc, const uint32_t* pkey, int pop_sp,
                                                      float* Chl, float* pop)
{
    if (st->err) return;
    __shared__ __align__(16) float buf[2][LCH];
    const PI_GLOBAL float* chl_store = gp(st->store[st->scur].Chl);
    const uint32_t nl = st->nlong;
    const uint32_t t = threadIdx.x;
    for (uint32_t b = K.b; b < nl; b += K.n) {
        const uint32_t c = lng[b];
        const uint32_t s0 = seg[c], s1 = seg[c + 1];
        float sum = 0.0f;
        stage_chl<4>(chl_store, psrc, s0, s1, t, 256, buf[0]);
        __syncthreads();
        uint32_t k = 0;
        for (uint32_t base = s0; base < s1; base += LCH, ++k) {
            const uint32_t nb = base + LCH;
            if (t >= 64) {
                if (nb < s1) stage_chl<6>(chl_store, psrc, nb, s1, t - 64, 192, buf[(k + 1) & 1]);
            } else if (t == 0) {
                sum = chain_sum(sum, buf[k & 1], min((uint32_t)LCH, s1 - base));
            }
            __syncthreads();
        }
        if (t == 0) {
            pop[c] = (float)(pop_sp < 0 ? s1 - s0 : species_count(pkey, s0, s1, (uint32_t)pop_sp));
            Chl[c] = sum;
        }
    }
}
__device__ __forceinline__ void k_cellsum_long_body(const BinArgs& x, Blk K) { k_cellsum_long_impl(K, x.st, x.lng, x.seg, x.psrc, x.pkey, x.pop_sp, x.Chl, x.pop); }
PI_KERNEL(k_cellsum_long, BinArgs, 256)

/* Long re-sorted segments, one block per cell.  The cell holds its stayers,
 * already in key order, then its arrivals in any order.  With at most MAXARR
 * arrivals: they are sorted in LDS (bitonic) and both runs are merged, each
 * entry's place = its rank in its own run + the number of keys of the other
 * run below it (keys are unique), written to the scratch region
 * [2*s0, 2*s0 + n) that no other segment overlaps and copied back.  With
 * more, the whole segment is sorted in the scratch region [2*s0, 2*s0 + pow2(n)). */
/* the number of keys below key in the sorted k[0..A2), A2 a power of two:
 * a fixed number of steps, no branch */
__device__ __forceinline__ uint32_t rank_below(const uint32_t* k, uint32_t A2, uint32_t key)
{
    uint32_t p = 0;
    for (uint32_t h = A2; h > 0; h >>= 1)
        if (p + h <= A2 && k[p + h - 1] < key) p += h;
    return p;
}
__device__ __forceinline__ uint32_t count_below(const uint32_t* k, uint32_t n, uint32_t key)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (k[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ void k_bigcells_impl(Blk K, DevState* st, const uint32_t* big, const uint32_t* seg,
                                                  const uint32_t* segp, const uint32_t* P,
                                                  uint32_t* pkey, uint32_t* psrc, uint32_t* skey, uint32_t* ssrc)
{
    if (st->err) return;
    __shared__ uint32_t AK[MAXARR], AS[MAXARR];
    const uint32_t nb = st->nbig;
    const bool have = st->nA > 0;
    for (uint32_t b = K.b; b < nb; b += K.n) {
        const uint32_t c = big[b];
        const uint32_t s0 = seg[c], n = seg[c + 1] - s0;
        const uint32_t nst = have ? P[segp[c + 1]] - P[segp[c]] : 0u;   /* stayers, as k_cellcount */
        const uint32_t m = n - nst;
        if (m <= MAXARR) {
            uint32_t A2 = 1;
            while (A2 < m) A2 <<= 1;
            for (uint32_t i = threadIdx.x; i < A2; i += blockDim.x) {
                AK[i] = i < m ? pkey[s0 + nst + i] : 0xFFFFFFFFu;
                AS[i] = i < m ? psrc[s0 + nst + i] : 0u;
            }
            __syncthreads();
            for (uint32_t k = 2; k <= A2; k <<= 1) {
                for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                    for (uint32_t i = threadIdx.x; i < A2; i += blockDim.x) {
                        const uint32_t l = i ^ j;
                        if (l > i) {
                            const bool up = (i & k) == 0;
                            const uint32_t ki = AK[i], kl = AK[l];
                            if ((ki > kl) == up) {
                                AK[i] = kl; AK[l] = ki;
                                const uint32_t t = AS[i]; AS[i] = AS[l]; AS[l] = t;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            uint32_t* K = skey + 2ull * s0;
            uint32_t* S = ssrc + 2ull * s0;
            for (uint32_t base = 0; base < nst; base += 4 * 256) {
                uint32_t key[4], src[4], pos[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t i = base + u * 256 + threadIdx.x;
                    key[u] = i < nst ? pkey[s0 + i] : 0u;
                    src[u] = i < nst ? psrc[s0 + i] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) pos[u] = base + u * 256 + threadIdx.x + rank_below(AK, A2, key[u]);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (base + u * 256 + threadIdx.x < nst) { K[pos[u]] = key[u]; S[pos[u]] = src[u]; }
                }
            }
            for (uint32_t base = 0; base < m; base += 4 * 256) {
                uint32_t pos[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t q = base + u * 256 + threadIdx.x;
                    pos[u] = q < m ? q + count_below(pkey + s0, nst, AK[q]) : 0u;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const uint32_t q = base + u * 256 + threadIdx.x;
                    if (q < m) { K[pos[u]] = AK[q]; S[pos[u]] = AS[q]; }
                }
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                pkey[s0 + i] = K[i];
                psrc[s0 + i] = S[i];
            }
            __syncthreads();
            continue;
        }
        uint32_t P2 = 1;
        while (P2 < n) P2 <<= 1;
        uint32_t* K = skey + 2ull * s0;
        uint32_t* S = ssrc + 2ull * s0;
        for (uint32_t i = threadIdx.x; i < P2; i += blockDim.x) {
            K[i] = i < n ? pkey[s0 + i] : 0xFFFFFFFFu;
            S[i] = i < n ? psrc[s0 + i] : 0u;
        }
        __syncthreads();
        for (uint32_t k = 2; k <= P2; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < P2; i += blockDim.x) {
                    const uint32_t l = i ^ j;
                    if (l > i) {
                        const bool up = (i & k) == 0;
                        const uint32_t ki = K[i], kl = K[l];
                        if ((ki > kl) == up) {
                            K[i] = kl; K[l] = ki;
                            const uint32_t t = S[i]; S[i] = S[l]; S[l] = t;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            pkey[s0 + i] = K[i];
            psrc[s0 + i] = S[i];
        }
        __syncthreads();
    }
}
__device__ __forceinline__ void k_bigcells_body(const BinArgs& x, Blk K) { k_bigcells_impl(K, x.st, x.big, x.seg, x.segp, x.P, x.pkey, x.psrc, x.skey, x.ssrc); }
PI_KERNEL(k_bigcells, BinArgs, 256)

/* calc_par_kernel! looped over levels (Individuals/utils.jl:55-68, time_step.jl:63-67) */
#define PAR_TC 64        /* columns per tile */
#define PAR_NZ 64        /* levels a tile holds in LDS (deeper grids take the column loop) */
__device__ __forceinline__ void k_par_impl(Blk K, DevState* st, DevGrid g, const float* Chl, const float* PARF_in,
                                             float* PARF_out, float* diag_par, float kc, float kw)
{
    if (st->err) return;
    PI_GLOBAL float* par = gp(st->par[1 - st->pcur]);
    const int ncol = g.Nx * g.Ny;
    if (g.Nz > PAR_NZ) {
        for (int col = K.b * blockDim.x + threadIdx.x; col < ncol; col += K.n * blockDim.x) {
            const int ii = col % g.Nx + 1 + g.Hx, jj = col / g.Nx + 1 + g.Hy;
            float P = PARF_in[col];
            for (int k = 1; k <= g.Nz; ++k) {
                const int kk = k + g.Hz;
                const long long c = gidx(g, ii, jj, kk);
                const float vol = cell_vol(g, c, kk);
                const float at = (Chl[c] / vol * kc + kw) * dzF_at(g, kk);
                const float ex = pi_expf(-at);
                const float p = P * (1.0f - ex) / at;
                par[c] = p;
                diag_par[c] = diag_par[c] + p;
                P = P * ex;
            }
            PARF_out[col] = P;
        }
        return;
    }
    /* Tiles of PAR_TC columns: the attenuations and their exponentials of every
     * (column, level) in parallel, then one lane per column runs the product
     * P·(1 − ex), P ← P·ex down the column (the only sequential part), then the
     * quotients by the attenuation and the stores in parallel.  Same operations
     * in the same order as the column loop above. */
    __shared__ float s_at[PAR_NZ * PAR_TC], s_q[PAR_NZ * PAR_TC];
    const int nitem = g.Nz * PAR_TC;
    constexpr int NR = PAR_NZ * PAR_TC / 256;          /* items per thread (blockDim.x = 256) */
    for (int col0 = K.b * PAR_TC; col0 < ncol; col0 += K.n * PAR_TC) {
        float dgv[NR];                                 /* diag_par of the items, loaded with the attenuations */
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int it = (int)threadIdx.x + r * 256;
            const int col = col0 + (it & (PAR_TC - 1)), kk = it / PAR_TC + 1 + g.Hz;
            dgv[r] = 0.0f;
            if (it >= nitem || col >= ncol) continue;
            const long long c = gidx(g, col % g.Nx + 1 + g.Hx, col / g.Nx + 1 + g.Hy, kk);
            const float vol = cell_vol(g, c, kk);
            const float at = (Chl[c] / vol * kc + kw) * dzF_at(g, kk);
            dgv[r] = diag_par[c];
            s_at[it] = at;
            s_q[it] = pi_expf(-at);
        }
        __syncthreads();
        if (threadIdx.x < PAR_TC && col0 + (int)threadIdx.x < ncol) {
            float P = PARF_in[col0 + threadIdx.x];
            for (int k = 0; k < g.Nz; ++k) {
                const float ex = s_q[k * PAR_TC + threadIdx.x];
                s_q[k * PAR_TC + threadIdx.x] = P * (1.0f - ex);
                P = P * ex;
            }
            PARF_out[col0 + threadIdx.x] = P;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int it = (int)threadIdx.x + r * 256;
            const int col = col0 + (it & (PAR_TC - 1)), kk = it / PAR_TC + 1 + g.Hz;
            if (it >= nitem || col >= ncol) continue;
            const long long c = gidx(g, col % g.Nx + 1 + g.Hx, col / g.Nx + 1 + g.Hy, kk);
            const float p = s_q[it] / s_at[it];
            par[c] = p;
            diag_par[c] = dgv[r] + p;
        }
        __syncthreads();
    }
}
__device__ __forceinline__ void k_par_body(const ParArgs& x, Blk K) { k_par_impl(K, x.st, x.g, x.Chl, x.PARF_in, x.PARF_out, x.diag_par, x.kc, x.kw); }
PI_KERNEL(k_par, ParArgs, 256)

/* ------------------------------------------------------------- physiology */
struct PState { float x, y, z, Sz, Bm, Cq, Nq, Pq, Chl, gen, age; float ex[PI_NEX]; };
struct Trs { float NH4, NO3, PO4, DOC, DFe, par, dpar, T, pop; };
struct Rates { float PS, VDOC, VNH4, VNO3, VPO4, resp, BS, exu, rho; };

/* shape_func_dec / shape_func_inc (Individuals/utils.jl:109-121) */
__device__ __forceinline__ float shape_dec(float x, float xmax, float k, int pw)
{
    float fx = dmaxf(0.0f, dminf(1.0f, 1.0f - x / xmax));
    float fp = pw == 2 ? pi_pow2f(fx) : pi_pow4f(fx);
    return fp / (k + fp);
}
__device__ __forceinline__ float shape_inc(float x, float xmax, float k, int pw)
{
    return 1.0f - shape_dec(x, xmax, k, pw);
}
/* tempFunc_PS / tempFunc (QuotaMode/growth_kernels.jl:1-17).  Both are pure
 * in (T, species) and share exp(-Ea / (R (T + 273.15))); OGT depends on the
 * species parameters only.  They are evaluated once per individual (TempF) and
 * reused by every call site of plankton_growth!, which is bit-identical. */
struct TempF { float ps, tf; };
__device__ __forceinline__ float ogt_of(const float* q)
{
    return pi_expf(-q[PI_P_Ea] / (8.3145f * (q[PI_P_Topt] + 273.15f)));
}
__device__ __forceinline__ TempF temp_funcs(float T, const float* q, float OGT)
{
    const float kx = pi_expf(-q[PI_P_Ea] / (8.3145f * (T + 273.15f)));
    TempF r;
    const float x = T - q[PI_P_Topt], xmax = q[PI_P_Tmax] - q[PI_P_Topt];
    const float regT = shape_dec(x, xmax, 4.0e-2f, 4);
    r.ps = dminf(1.0f, dmaxf(0.0f, kx * regT) / OGT);
    r.tf = dminf(1.0f, dmaxf(0.0f, kx) / OGT);
    return r;
}

/* plankton_growth! for one individual (QuotaMode/plankton_growth.jl:2-24,
 * growth_kernels.jl:48-171, division_death_probability.jl:176-278). */
__device__ void growth(PState& s, const Trs& t, const float* q, float OGT, bool ac, float dt, float tmodel,
                       int dvid_type, float U0, float U1, float U2, Rates& r,
                       bool& dgraz, bool& dmort, bool& ddvid)
{
    const float e30 = 1.0e-30f;
    const TempF T = temp_funcs(t.T, q, OGT);
    {   /* calc_inorganic_uptake_kernel! */
        float aI = t.par * q[PI_P_alpha] * q[PI_P_Phi];
        float PCm = q[PI_P_PCmax] * T.ps;
        float PS = PCm * (1.0f - pi_expf(-aI / dmaxf(e30, PCm) * s.Chl / dmaxf(e30, s.Bm))) * s.Bm;
        r.PS = jl_maskf(ac, PS);
        float Qn = (s.Nq + s.Bm * q[PI_P_R_NC]) / dmaxf(e30, s.Bm + s.Cq);
        float Qp = (s.Pq + s.Bm * q[PI_P_R_PC]) / dmaxf(e30, s.Bm + s.Cq);
        float regQN = shape_dec(Qn, q[PI_P_Nqmax], 1.0e-4f, 4);
        float regQP = shape_dec(Qp, q[PI_P_Pqmax], 1.0e-4f, 4);
        float tf = T.tf;
        float VNH4 = jl_maskf(ac, q[PI_P_VNH4max] * regQN * t.NH4 / dmaxf(e30, t.NH4 + q[PI_P_KsatNH4]) * tf * s.Bm);
        float VNO3 = jl_maskf(ac, q[PI_P_VNO3max] * regQN * t.NO3 / dmaxf(e30, t.NO3 + q[PI_P_KsatNO3]) * tf * s.Bm);
        float VPO4 = jl_maskf(ac, q[PI_P_VPO4max] * regQP * t.PO4 / dmaxf(e30, t.PO4 + q[PI_P_KsatPO4]) * tf * s.Bm);
        float pp = dmaxf(1.0f, t.pop);
        r.VNH4 = dminf(VNH4, t.NH4 / dt / pp);
        r.VNO3 = dminf(VNO3, t.NO3 / dt / pp);
        r.VPO4 = dminf(VPO4, t.PO4 / dt / pp);
    }
    /* update_quotas_1_kernel! */
    s.Cq = s.Cq + r.PS * dt;
    s.Pq = s.Pq + r.VPO4 * dt;
    s.Nq = s.Nq + (r.VNH4 + r.VNO3) * dt;
    {   /* calc_organic_uptake_kernel! */
        float Qc = s.Cq / dmaxf(e30, s.Bm + s.Cq);
        float regQ = shape_dec(Qc, q[PI_P_Cqmax], 1.0e-4f, 4);
        float VN = q[PI_P_VDOCmax] * regQ * t.DOC / dmaxf(e30, t.DOC + q[PI_P_KsatDOC]) * T.tf * s.Bm;
        float VD = dminf(VN, t.DOC / dt / dmaxf(1.0f, t.pop));
        r.VDOC = jl_maskf(ac, VD);
        r.VDOC = jl_maskf(jl_islessf(1.0e-2f, r.PS / (r.VDOC + r.PS)), r.VDOC);
    }
    {   /* calc_ρChl_kernel! */
        float rr = r.PS / dmaxf(e30, s.Bm) * q[PI_P_Chl2N] /
                   dmaxf(e30, t.par * q[PI_P_alpha] * q[PI_P_Phi] * s.Chl / s.Bm);
        rr = jl_maskf(jl_islessf(1.0e-1f, t.par), rr);
        r.rho = jl_maskf(ac, rr);
    }
    /* calc_respir_kernel! */
    r.resp = jl_maskf(ac, q[PI_P_respir] * s.Bm * T.tf);
    /* update_quotas_2_kernel! */
    s.Cq = s.Cq + (r.VDOC - r.resp) * dt;
    s.Bm = s.Bm - dmaxf(0.0f, (0.0f - s.Cq));
    s.Nq = s.Nq + dmaxf(0.0f, (0.0f - s.Cq)) * q[PI_P_R_NC];
    s.Pq = s.Pq + dmaxf(0.0f, (0.0f - s.Cq)) * q[PI_P_R_PC];
    s.Cq = s.Cq + dmaxf(0.0f, (0.0f - s.Cq));
    {   /* calc_BS_kernel! */
        float mn = dminf(dminf(s.Cq, s.Nq / q[PI_P_R_NC]), s.Pq / q[PI_P_R_PC]);
        r.BS = jl_maskf(ac, mn * q[PI_P_k_mtb]);
        r.exu = jl_maskf(ac, dmaxf(0.0f, s.Cq - mn) * q[PI_P_k_mtb]);
    }
    /* update_biomass_kernel! */
    s.Bm = s.Bm + dt * r.BS;
    s.Cq = s.Cq - dt * (r.BS + r.exu);
    s.Nq = s.Nq - dt * r.BS * q[PI_P_R_NC];
    s.Pq = s.Pq - dt * r.BS * q[PI_P_R_PC];
    s.Chl = s.Chl + dt * r.BS * q[PI_P_R_NC] * r.rho;
    s.age = s.age + jl_maskf(ac, dt / 3600.0f);
    /* update_cellsize_kernel! */
    s.Sz = s.Bm / (q[PI_P_Cquota] * q[PI_P_Nsuper]);
    /* probabilities + get_probability! */
    float graz = jl_maskf(ac, t.pop * q[PI_P_grz_P]);
    float mort = jl_maskf(ac, q[PI_P_mort_P] * shape_dec(s.Sz, q[PI_P_mort_reg], 1.0e-5f, 4));
    float dv;
    const float hour = fmodf(tmodel, 86400.0f) / 3600.0f;
    if (dvid_type == 1) dv = q[PI_P_dvid_P] * shape_inc(s.Sz, q[PI_P_dvid_reg], 2.0e-3f, 2);
    else if (dvid_type == 2) dv = q[PI_P_dvid_P] * shape_inc(hour, q[PI_P_dvid_reg], 1.0e-5f, 4);
    else dv = q[PI_P_dvid_P] * shape_inc(s.Sz, q[PI_P_dvid_reg], 1.0e-3f, 4) * shape_inc(hour, q[PI_P_dvid_reg2], 1.0e-5f, 4);
    float dvid = jl_maskf(ac, dv);
    dgraz = jl_islessf(U0, graz * dt) && ac;
    dmort = jl_islessf(U1, mort * dt) && ac;
    ddvid = jl_islessf(U2, dvid * dt) && ac;
}

/* CarbonMode plankton_growth! for one individual (CarbonMode/plankton_growth.jl:2-29,
 * growth_kernels.jl:1-120, division_death_probability.jl:56-136).  s.Cq holds
 * the damaged biomass Bd; the rates returned are PS and resp (RS). */
__device__ __forceinline__ float ogt_carbon(const float* q)
{
    return pi_expf(-q[PI_C_Ea] / (8.3145f * (q[PI_C_Topt] + 273.15f)));
}
__device__ void growth_carbon(PState& s, const Trs& t, const float* q, float OGT, bool ac, float dt, float tmodel,
                              int dvid_type, float U0, float U1, float U2, Rates& r,
                              bool& dgraz, bool& dmort, bool& ddvid)
{
    const float e30 = 1.0e-30f;
    const bool thermal = q[PI_C_thermal] == 1.0f;
    /* tempFunc / tempFunc_PS (growth_kernels.jl:1-17), regT with k = 0.1 */
    const float kx = pi_expf(-q[PI_C_Ea] / (8.3145f * (t.T + 273.15f)));
    const float x = t.T - q[PI_C_Topt], xmax = q[PI_C_Tmax] - q[PI_C_Topt];
    const float tps = dminf(1.0f, dmaxf(0.0f, kx * shape_dec(x, xmax, 1.0e-1f, 4)) / OGT);
    const float tf = dminf(1.0f, dmaxf(0.0f, kx) / OGT);
    /* calc_PS_kernel! (calc_photosynthesis) */
    const float aI = t.par * q[PI_C_alpha] * q[PI_C_Phi];
    const float PCm = q[PI_C_PCmax] * (tf * q[PI_C_thermal] + tps * (1.0f - q[PI_C_thermal]));
    const float light = 1.0f - pi_expf(-aI / dmaxf(e30, PCm) * s.Chl / dmaxf(e30, s.Bm));
    const float PS = jl_maskf(ac, PCm * dmaxf(0.0f, s.Bm - s.Cq) *
                                      (light * (1.0f - q[PI_C_is_bact]) + q[PI_C_is_bact]));
    /* gamma_alloc, calc_BS_kernel!, calc_repair_kernel! */
    float gam = jl_maskf(jl_islessf(0.0f, s.Cq), s.Cq / dmaxf(e30, s.Bm));
    gam = jl_maskf(thermal, dminf(1.0f, gam));
    const float BS = PS * (1.0f - gam);
    const float RP = PS * gam;
    /* calc_thermal_damage_kernel! (not masked by ac) */
    float TD = (t.T - q[PI_C_Topt]) * q[PI_C_f_T2B] * (s.Bm - s.Cq);
    TD = jl_maskf(jl_islessf(q[PI_C_Topt], t.T), TD);
    TD = jl_maskf(jl_islessf(s.Cq, s.Bm), TD);
    TD = jl_maskf(thermal, TD);
    TD = dmaxf(0.0f, dminf(TD, (s.Bm - s.Cq) / dt));
    /* calc_respir_kernel! */
    const float RS = jl_maskf(ac, q[PI_C_respir] * s.Bm * tf);
    /* update_quotas_kernel!, update_cellsize_kernel! */
    s.Bm = s.Bm + (BS - RS) * dt;
    s.Cq = s.Cq + (TD - RP) * dt;
    s.age = s.age + jl_maskf(ac, dt / 3600.0f);
    s.Sz = s.Bm / (q[PI_C_Cquota] * q[PI_C_Nsuper]);
    s.Chl = s.Bm * q[PI_C_Chl2C];
    r.PS = PS; r.resp = RS;
    r.VDOC = r.VNH4 = r.VNO3 = r.VPO4 = r.BS = r.exu = r.rho = 0.0f;
    /* probabilities + get_probability! */
    const float graz = jl_maskf(ac, t.pop * q[PI_C_grz_P]);
    float dv;
    const float hour = fmodf(tmodel, 86400.0f) / 3600.0f;
    if (dvid_type == 1) dv = q[PI_C_dvid_P] * shape_inc(s.Sz, q[PI_C_dvid_reg], 2.0e-3f, 2);
    else if (dvid_type == 2) dv = q[PI_C_dvid_P] * shape_inc(hour, q[PI_C_dvid_reg], 1.0e-5f, 4);
    else dv = q[PI_C_dvid_P] * shape_inc(s.Sz, q[PI_C_dvid_reg], 1.0e-3f, 4) * shape_inc(hour, q[PI_C_dvid_reg2], 1.0e-5f, 4);
    const float dvid = jl_maskf(ac, dv);
    const float mort = thermal
        ? jl_maskf(ac, q[PI_C_mort_P] * shape_inc(s.Cq / dmaxf(e30, s.Bm), 0.99f, 1.0e-6f, 4))
        : jl_maskf(ac, q[PI_C_mort_P] * shape_dec(s.Sz, q[PI_C_mort_reg], 1.0e-5f, 4));
    dgraz = jl_islessf(U0, graz * dt) && ac;
    dmort = jl_islessf(U1, mort * dt) && ac;
    ddvid = jl_islessf(U2, dvid * dt) && ac;
}

/* MacroMolecularMode plankton_growth! for one individual
 * (MacroMolecularMode/plankton_growth.jl:2-28, growth_kernels.jl:1-200,
 * division_death_probability.jl:39-44,96-101,258-278).  Columns: Bm = PRO,
 * Sz = DNA, ex[0] = RNA, Cq = CH, Nq = NST, Pq = PST.  The rates returned are
 * PS, VDOC, VNH4, VNO3, VPO4, resp and exu (calc_consume!). */
__device__ __forceinline__ float ogt_mm(const float* q)
{
    return pi_expf(-q[PI_M_Ea] / (8.3145f * (q[PI_M_Topt] + 273.15f)));
}
/* total_C_biomass / total_N_biomass / total_P_biomass (plankton_generation.jl:84-95) */
__device__ __forceinline__ float mm_totC(const PState& s)
{
    return s.Bm + s.Sz + s.ex[0] + s.Cq + s.Chl / 893.49f * 55.0f;
}
__device__ __forceinline__ float mm_totN(const PState& s, const float* q)
{
    return s.Bm * q[PI_M_R_NC_PRO] + s.Sz * q[PI_M_R_NC_DNA] + s.ex[0] * q[PI_M_R_NC_RNA] + s.Nq + s.Chl / 893.49f * 4.0f;
}
__device__ __forceinline__ float mm_totP(const PState& s, const float* q)
{
    return s.Sz * q[PI_M_R_PC_DNA] + s.ex[0] * q[PI_M_R_PC_RNA] + s.Pq;
}
__device__ void growth_mm(PState& s, const Trs& t, const float* q, float OGT, bool ac, float dt,
                          float U0, float U1, float U2, Rates& r, bool& dgraz, bool& dmort, bool& ddvid)
{
    const float e30 = 1.0e-30f;
    /* tempFunc_PS / tempFunc (growth_kernels.jl:1-17) */
    const float kx = pi_expf(-q[PI_M_Ea] / (8.3145f * (t.T + 273.15f)));
    const float x = t.T - q[PI_M_Topt], xmax = q[PI_M_Tmax] - q[PI_M_Topt];
    const float tps = dminf(1.0f, dmaxf(0.0f, kx * shape_dec(x, xmax, 4.0e-2f, 4)) / OGT);
    const float tf = dminf(1.0f, dmaxf(0.0f, kx) / OGT);
    {   /* calc_inorganic_uptake_kernel! (calc_PS, calc_NP_uptake) */
        const float aI = t.par * q[PI_M_alpha] * q[PI_M_Phi];
        const float PCm = q[PI_M_PCmax] * tps;
        r.PS = jl_maskf(ac, PCm * (1.0f - pi_expf(-aI / dmaxf(e30, PCm) * s.Chl / dmaxf(e30, s.Bm))) * s.Bm);
        const float Rn = s.Nq / dmaxf(e30, mm_totN(s, q));
        const float Rp = s.Pq / dmaxf(e30, mm_totP(s, q));
        const float regQN = shape_dec(Rn, q[PI_M_NSTmax], 1.0e-4f, 4);
        const float regQP = shape_dec(Rp, q[PI_M_PSTmax], 1.0e-4f, 4);
        const float VNH4 = jl_maskf(ac, q[PI_M_VNH4max] * regQN * t.NH4 / dmaxf(e30, t.NH4 + q[PI_M_KsatNH4]) * tf * s.Bm);
        const float VNO3 = jl_maskf(ac, q[PI_M_VNO3max] * regQN * t.NO3 / dmaxf(e30, t.NO3 + q[PI_M_KsatNO3]) * tf * s.Bm);
        const float VPO4 = jl_maskf(ac, q[PI_M_VPO4max] * regQP * t.PO4 / dmaxf(e30, t.PO4 + q[PI_M_KsatPO4]) * tf * s.Bm);
        const float pp = dmaxf(1.0f, t.pop);
        r.VNH4 = dminf(VNH4, t.NH4 / dt / pp);
        r.VNO3 = dminf(VNO3, t.NO3 / dt / pp);
        r.VPO4 = dminf(VPO4, t.PO4 / dt / pp);
    }
    /* update_quotas_1_kernel! */
    s.Cq = s.Cq + r.PS * dt;
    s.Pq = s.Pq + r.VPO4 * dt;
    s.Nq = s.Nq + (r.VNH4 + r.VNO3) * dt;
    {   /* calc_organic_uptake_kernel! (calc_DOC_uptake) */
        const float regQ = shape_dec(s.Cq / dmaxf(e30, mm_totC(s)), q[PI_M_CHmax], 1.0e-4f, 4);
        const float VN = q[PI_M_VDOCmax] * regQ * t.DOC / dmaxf(e30, t.DOC + q[PI_M_KsatDOC]) * tf * s.Bm;
        r.VDOC = jl_maskf(ac, dminf(VN, t.DOC / dt / dmaxf(1.0f, t.pop)));
        r.VDOC = jl_maskf(jl_islessf(0.05f, r.PS / (r.VDOC + r.PS)), r.VDOC);
    }
    {   /* calc_ρChl_kernel! */
        float rr = r.PS / dmaxf(e30, s.Bm) / dmaxf(e30, t.par * q[PI_M_alpha] * q[PI_M_Phi] * s.Chl / dmaxf(e30, s.Bm));
        rr = jl_maskf(jl_islessf(0.1f, t.par), rr);
        r.rho = jl_maskf(ac, rr);
    }
    /* calc_respir_kernel! */
    r.resp = jl_maskf(ac, q[PI_M_respir] * s.Bm * tf);
    /* update_quotas_2_kernel!: respiration beyond CH is taken from protein */
    s.Cq = s.Cq + (r.VDOC - r.resp) * dt;
    s.Bm = s.Bm - dmaxf(0.0f, (0.0f - s.Cq));
    s.Nq = s.Nq + dmaxf(0.0f, (0.0f - s.Cq)) * q[PI_M_R_NC_PRO];
    s.Cq = s.Cq + dmaxf(0.0f, (0.0f - s.Cq));
    /* calc_BS_kernel! (not masked by ac) */
    const float kN = q[PI_M_Nsuper];
    const float lPRO = dminf(s.Cq / (s.Cq + q[PI_M_k_sat_pro] * kN),
                             s.Nq / (s.Nq + q[PI_M_k_sat_pro] * kN * q[PI_M_R_NC_PRO]));
    const float lDNA = dminf(dminf(s.Cq / (s.Cq + q[PI_M_k_sat_dna] * kN),
                                   s.Nq / (s.Nq + q[PI_M_k_sat_dna] * kN * q[PI_M_R_NC_DNA])),
                             s.Pq / (s.Pq + q[PI_M_k_sat_dna] * kN * q[PI_M_R_PC_DNA]));
    const float lRNA = dminf(dminf(s.Cq / (s.Cq + q[PI_M_k_sat_rna] * kN),
                                   s.Nq / (s.Nq + q[PI_M_k_sat_rna] * kN * q[PI_M_R_NC_RNA])),
                             s.Pq / (s.Pq + q[PI_M_k_sat_rna] * kN * q[PI_M_R_PC_RNA]));
    const float SPRO = q[PI_M_k_pro] * s.ex[0] * lPRO * tf;
    const float SDNA = jl_maskf(jl_islessf(s.Sz / (q[PI_M_C_DNA] * kN), 2.0f), q[PI_M_k_dna] * s.Bm * lDNA * tf);
    const float SRNA = q[PI_M_k_rna] * s.Bm * lRNA * tf;
    /* update_biomass_kernel! */
    s.Bm = s.Bm + dt * SPRO;
    s.Sz = s.Sz + dt * SDNA;
    s.ex[0] = s.ex[0] + dt * SRNA;
    s.Cq = s.Cq - dt * (SPRO + SDNA + SRNA + SPRO * r.rho);
    s.Nq = s.Nq - dt * (SPRO * q[PI_M_R_NC_PRO] + SDNA * q[PI_M_R_NC_DNA] + SRNA * q[PI_M_R_NC_RNA] +
                        SPRO * r.rho * 4.0f / 55.0f);
    s.Pq = s.Pq - dt * (SDNA * q[PI_M_R_PC_DNA] + SRNA * q[PI_M_R_PC_RNA]);
    s.Chl = s.Chl + dt * SPRO * r.rho * 893.49f / 55.0f;
    s.age = s.age + jl_maskf(ac, dt / 3600.0f);
    /* calc_exudation_kernel!, update_CH_kernel! (not masked by ac) */
    r.exu = dmaxf(0.0f, s.Cq - q[PI_M_CHmax] * mm_totC(s));
    s.Cq = s.Cq - r.exu;
    r.BS = 0.0f;
    /* calc_graz_quadratic!, calc_MM_mort!, calc_MM_dvid!, get_probability! */
    const float graz = jl_maskf(ac, t.pop * q[PI_M_grz_P]);
    const float mort = jl_maskf(ac, q[PI_M_mort_P] * shape_inc(s.age, q[PI_M_mort_reg], 1.0e-4f, 4));
    const float dvid = jl_maskf(jl_islessf(2.0f, s.Sz / (q[PI_M_C_DNA] * q[PI_M_Nsuper])), q[PI_M_dvid_P]);
    dgraz = jl_islessf(U0, graz * dt) && ac;
    dmort = jl_islessf(U1, mort * dt) && ac;
    ddvid = jl_islessf(U2, dvid * dt) && ac;
}

/* IronEnergyMode plankton_growth! for one individual (IronEnergyMode/plankton_growth.jl:2-30,
 * growth_kernels.jl:1-366, division_death_probability.jl).  Columns: Sz, Bm, Chl;
 * Cq = CH, Nq = qNH4, Pq = qP; ex = qNO3, qFe, qFePS, qFeNR, qFeNF, tdark, exEn,
 * ptc.  The rates returned for calc_consume! are PS, resp (= RS after the energy
 * allocation), VNH4, VNO3, VPO4 and VFe (in VDOC's place). */
__device__ __forceinline__ float ogt_iron(const float* q)
{
    return pi_expf(-q[PI_I_Ea] / (8.3145f * (q[PI_I_Topt] + 273.15f)));
}
__device__ void growth_iron(PState& s, const Trs& t, const float* q, float OGT, bool ac, float dt, float tmodel,
                            int dvid_type, float U0, float U1, float U2, Rates& r,
                            bool& dgraz, bool& dmort, bool& ddvid)
{
    const float e30 = 1.0e-30f;
    float &CH = s.Cq, &qNH4 = s.Nq, &qP = s.Pq;
    float &qNO3 = s.ex[0], &qFe = s.ex[1], &qFePS = s.ex[2], &qFeNR = s.ex[3], &qFeNF = s.ex[4];
    float &tdark = s.ex[5], &exEn = s.ex[6];
    const float is_nr = q[PI_I_is_nr], is_nf = q[PI_I_is_croc] + q[PI_I_is_tric];
    /* tempFunc_CF / tempFunc (growth_kernels.jl:1-17) */
    const float kx = pi_expf(-q[PI_I_Ea] / (8.3145f * (t.T + 273.15f)));
    const float xT = t.T - q[PI_I_Topt], xmax = q[PI_I_Tmax] - q[PI_I_Topt];
    const float tcf = dminf(1.0f, dmaxf(0.0f, kx * shape_dec(xT, xmax, 4.0e-2f, 4)) / OGT);
    const float tf = dminf(1.0f, dmaxf(0.0f, kx) / OGT);
    const float pp = dmaxf(1.0f, t.pop);
    float VFe;
    {   /* calc_trs_uptake_kernel! (calc_NH4_uptake, calc_NO3_uptake, calc_P_uptake, calc_Fe_uptake) */
        const float BC = dmaxf(e30, s.Bm + CH);
        const float rNH4 = shape_dec(qNH4 / BC, q[PI_I_qNH4max], 1.0e-4f, 2);
        const float VNH4 = jl_maskf(ac, q[PI_I_VNH4max] * rNH4 * t.NH4 / dmaxf(e30, t.NH4 + q[PI_I_KsatNH4]) * tf * s.Bm);
        r.VNH4 = dminf(VNH4, t.NH4 / dt / pp) * is_nr;
        const float rNO3 = shape_dec(qNO3 / BC, q[PI_I_qNO3max], 1.0e-4f, 2);
        const float VNO3 = jl_maskf(ac, q[PI_I_VNO3max] * rNO3 * t.NO3 / dmaxf(e30, t.NO3 + q[PI_I_KsatNO3]) * tf * s.Bm);
        r.VNO3 = dminf(VNO3, t.NO3 / dt / pp) * is_nr;
        const float rP = shape_dec((qP + s.Bm * q[PI_I_R_PC]) / BC, q[PI_I_qPmax], 1.0e-4f, 2);
        const float VPO4 = jl_maskf(ac, q[PI_I_VPO4max] * rP * t.PO4 / dmaxf(e30, t.PO4 + q[PI_I_KsatPO4]) * tf * s.Bm);
        r.VPO4 = dminf(VPO4, t.PO4 / dt / pp);
        const float rFe = shape_dec((qFe + qFePS + qFeNF + qFeNR) / BC, q[PI_I_qFemax], 1.0e-4f, 2);
        const float SA = q[PI_I_SA] * pi_powf(s.Sz, 2.0f / 3.0f);
        const float V = jl_maskf(ac, q[PI_I_KSAFe] * SA * t.DFe * rFe * q[PI_I_Nsuper] * tf);
        VFe = dminf(V, t.DFe / dt / pp);
    }
    /* update_state_1_kernel! */
    qP = qP + r.VPO4 * dt;
    qNO3 = qNO3 + r.VNO3 * dt;
    qNH4 = qNH4 + r.VNH4 * dt;
    qFe = qFe + VFe * dt;
    /* calc_PS_kernel! (not masked by ac) */
    float PS;
    {
        const float aI = t.par * q[PI_I_alpha] * s.Chl / dmaxf(e30, s.Bm);
        const float Qps = qFePS / dmaxf(e30, s.Bm + CH);
        const float Ksat = Qps / dmaxf(e30, Qps + q[PI_I_KfePS]);
        PS = q[PI_I_PCmax] * (1.0f - pi_expf(-aI)) * s.Bm * Ksat;
    }
    /* calc_respiration_kernel! */
    float RS = jl_maskf(ac, CH * q[PI_I_k_rs] * tf);
    RS = dminf(RS, CH / dt);
    float ERS = jl_maskf(ac, RS * q[PI_I_e_rs]);
    /* calc_carbon_fixation_kernel! */
    const float regQC = shape_dec(CH / dmaxf(e30, s.Bm + CH), q[PI_I_CHmax], 1.0e-4f, 2);
    float CF = jl_maskf(ac, q[PI_I_k_cf] * regQC * tcf * s.Bm);
    float ECF = jl_maskf(ac, CF * q[PI_I_e_cf]);
    /* calc_NO3_reduction_kernel! */
    const float regN = shape_dec(qNH4 / dmaxf(e30, s.Bm + CH), q[PI_I_qNH4max], 1.0e-4f, 2);
    float NR, ENR;
    {
        const float Qnr = qFeNR / dmaxf(e30, s.Bm + CH);
        const float Ksat = Qnr / dmaxf(e30, Qnr + q[PI_I_KfeNR]);
        float n = jl_maskf(ac, q[PI_I_k_nr] * regN * qNO3 * Ksat * tf);
        n = dminf(n, qNO3 / dt);
        const float en = jl_maskf(ac, n * q[PI_I_e_nr]);
        NR = n * is_nr;
        ENR = en * is_nr;
    }
    /* calc_nitrogen_fixation_kernel! */
    float NF, ENF;
    {
        const float Qnf = qFeNF / dmaxf(e30, s.Bm + CH);
        const float Ksat = Qnf / dmaxf(e30, Qnf + q[PI_I_KfeNF]);
        const float n = jl_maskf(ac, q[PI_I_k_nf] * regN * Ksat * tf * s.Bm);
        const float en = jl_maskf(ac, n * q[PI_I_e_nf]);
        NF = n * is_nf;
        ENF = en * is_nf;
    }
    /* energy_allocation_kernel! (energy_alloc) */
    {
        const float ECFt = dminf(PS, ECF);
        const float ENFt = dminf(ENF, PS + ERS - ECFt) * is_nf;
        const float ENRt = dminf(ENR, PS + ERS - ECFt) * is_nr;
        const float ERSt = dminf(ERS, dmaxf(0.0f, ECFt + ENFt + ENRt - PS));
        exEn = dmaxf(0.0f, PS - ECFt - ENFt - ENRt);
        ERS = ERSt; ECF = ECFt;
        RS = ERS / q[PI_I_e_rs];
        CF = ECF / q[PI_I_e_cf];
        NF = ENFt / q[PI_I_e_nf];
        NR = ENRt / q[PI_I_e_nr];
    }
    /* update_state_2_kernel! */
    CH = CH + (CF - RS) * dt;
    qNO3 = qNO3 + -NR * dt;
    qNH4 = qNH4 + NR * dt;
    qNH4 = qNH4 + NF * dt;
    /* calc_ρChl_kernel! */
    float rho = q[PI_I_Chl2N] / dmaxf(e30, t.par * q[PI_I_alpha] * s.Chl / s.Bm);
    rho = jl_maskf(ac, jl_maskf(jl_islessf(1.0e-1f, t.par), rho));
    /* calc_BS_kernel! */
    const float BS = jl_maskf(ac, dminf(dminf(CH, qNH4 / q[PI_I_R_NC]), qP / q[PI_I_R_PC]) * q[PI_I_k_mtb]);
    /* calc_iron_fluxes_kernel! (iron_alloc, not masked by ac); Bool sums are Int */
    float ST2PS, PS2ST, ST2NR, NR2ST, ST2NF, NF2ST;
    {
        const bool dark = jl_isequalf(0.0f, t.par), light = jl_islessf(0.0f, t.par);
        const bool rising = jl_islessf(0.0f, t.dpar), falling = jl_islessf(t.dpar, 0.0f);
        const bool early = jl_islessf(tdark, q[PI_I_NF_clock]), late = jl_islessf(q[PI_I_NF_clock], tdark);
        ST2PS = jl_maskf(rising, q[PI_I_k_Fe_ST2PS] * qFe);
        PS2ST = q[PI_I_k_Fe_PS2ST] * qFePS * (float)((int)falling + (int)dark);
        ST2NR = jl_maskf(early, jl_maskf(dark, q[PI_I_k_Fe_ST2NR] * qFe)) * q[PI_I_is_nr];
        NR2ST = q[PI_I_k_Fe_NR2ST] * qFeNR * (float)((int)(dark && late) + (int)light) * q[PI_I_is_nr];
        const float ST2NF_cr = jl_maskf(early, jl_maskf(dark, q[PI_I_k_Fe_ST2NF] * qFe));
        const float NF2ST_cr = q[PI_I_k_Fe_NF2ST] * qFeNF * (float)((int)(dark && late) + (int)light);
        const float ST2NF_tr = jl_maskf(rising, q[PI_I_k_Fe_ST2NF] * qFe);
        const float NF2ST_tr = q[PI_I_k_Fe_NF2ST] * qFeNF * (float)((int)falling + (int)dark);
        ST2NF = ST2NF_cr * q[PI_I_is_croc] + ST2NF_tr * q[PI_I_is_tric];
        NF2ST = NF2ST_cr * q[PI_I_is_croc] + NF2ST_tr * q[PI_I_is_tric];
        ST2PS = dminf(ST2PS, qFe / dt);
        PS2ST = dminf(PS2ST, qFePS / dt);
        ST2NR = dminf(ST2NR, qFe / dt - ST2PS);
        NR2ST = dminf(NR2ST, qFeNR / dt);
        ST2NF = dminf(ST2NF, qFe / dt - ST2PS);
        NF2ST = dminf(NF2ST, qFeNF / dt);
    }
    /* update_biomass_kernel! */
    s.Bm = s.Bm + dt * BS;
    CH = CH - dt * BS;
    qNH4 = qNH4 - dt * BS * q[PI_I_R_NC];
    qP = qP - dt * BS * q[PI_I_R_PC];
    s.Chl = s.Chl + dt * BS * q[PI_I_R_NC] * rho;
    s.age = s.age + jl_maskf(ac, dt / 3600.0f);
    qFe = qFe + dt * (PS2ST - ST2PS + NR2ST - ST2NR + NF2ST - ST2NF);
    qFePS = qFePS + dt * (ST2PS - PS2ST);
    qFeNR = qFeNR + dt * (ST2NR - NR2ST);
    qFeNF = qFeNF + dt * (ST2NF - NF2ST);
    /* update_cellsize_kernel!, update_tdark_kernel! */
    s.Sz = s.Bm / (q[PI_I_Cquota] * q[PI_I_Nsuper]);
    tdark = tdark * (1.0f - (float)jl_islessf(0.0f, t.par)) + jl_maskf(jl_isequalf(0.0f, t.par), dt);
    r.PS = PS; r.resp = RS; r.VDOC = VFe; r.BS = BS; r.exu = 0.0f; r.rho = rho;
    /* calc_graz_quadratic!, calc_mort!, calc_dvid!, get_probability! */
    const float graz = jl_maskf(ac, t.pop * q[PI_I_grz_P]);
    const float mort = jl_maskf(ac, q[PI_I_mort_P] * shape_dec(s.Sz, q[PI_I_mort_reg], 1.0e-5f, 4));
    float dv;
    const float hour = fmodf(tmodel, 86400.0f) / 3600.0f;
    if (dvid_type == 1) dv = q[PI_I_dvid_P] * shape_inc(s.Sz, q[PI_I_dvid_reg], 2.0e-3f, 2);
    else if (dvid_type == 2) dv = q[PI_I_dvid_P] * shape_inc(hour, q[PI_I_dvid_reg], 1.0e-5f, 4);
    else dv = q[PI_I_dvid_P] * shape_inc(s.Sz, q[PI_I_dvid_reg], 1.0e-3f, 4) * shape_inc(hour, q[PI_I_dvid_reg2], 1.0e-5f, 4);
    const float dvid = jl_maskf(ac, dv);
    dgraz = jl_islessf(U0, graz * dt) && ac;
    dmort = jl_islessf(U1, mort * dt) && ac;
    ddvid = jl_islessf(U2, dvid * dt) && ac;
}

/* k-th set bit (0-based) of a word */
__device__ __forceinline__ int select_bit(uint32_t w, uint32_t k)
{
    for (int b = 0; b < 32; ++b) {
        if (w & (1u << b)) {
            if (k == 0) return b;
            --k;
        }
    }
    return -1;
}


__device__ __forceinline__ uint32_t prefG(const PhysArgs& a, int which, int sp, uint32_t w)
{
    const uint32_t base = (uint32_t)which * (uint32_t)a.nsp * (uint32_t)a.nw + (uint32_t)sp * (uint32_t)a.nw;
    return a.G[base + w] - a.G[base];
}
__device__ __forceinline__ uint32_t valid_mask(const PhysArgs& a, uint32_t w)
{
    const long long lo = (long long)w * 32;
    const long long rem = a.cap - lo;
    return rem >= 32 ? 0xFFFFFFFFu : ((1u << rem) - 1u);
}

template <int MODE>
__device__ __forceinline__ void k_phys_body(const PhysArgs& a, Blk K)
{
    constexpr bool CARBON = MODE == PI_MODE_CARBON;
    constexpr bool MM = MODE == PI_MODE_MACROMOLECULAR;
    constexpr bool IRON = MODE == PI_MODE_IRONENERGY;
    constexpr int NEX = mode_nex(MODE);
    /* the extra columns divide_to_half! halves: RNA; the iron pools, qNO3 and
     * exEn (not tdark or ptc) */
    constexpr uint32_t HALF = MM ? 0x1u : IRON ? 0x5Fu : 0u;
    const DevState* st = a.st;
    if (st->err) return;
    const DevGrid& g = a.g;
    const uint32_t n = a.seg[g.ncell];
    const float dt = a.dt;
    const GStore IN = gstore(st->store[st->scur]), OUT = gstore(st->store[1 - st->scur]);
    const long long iteration = st->iteration + 1;
    const float tmodel = (float)iteration * dt;
    /* device-state values are read before the loop (see move_entry) */
    const PI_GLOBAL float* TRs = gp(st->tr[st->tcur][0]);
    const PI_GLOBAL float* TR_NH4 = TRs + (long long)PI_NH4 * g.ncell, *TR_NO3 = TRs + (long long)PI_NO3 * g.ncell;
    const PI_GLOBAL float* TR_PO4 = TRs + (long long)PI_PO4 * g.ncell, *TR_DOC = TRs + (long long)PI_DOC * g.ncell;
    const PI_GLOBAL float* TR_DFe = TRs + (long long)PI_DFe * g.ncell;
    const PI_GLOBAL float* PAR = gp(st->par[1 - st->pcur]);
    const PI_GLOBAL float* PAR0 = gp(st->par[st->pcur]);
    /* per-species tables in LDS: the parameters (read by every individual, with
     * per-lane species) and OGT, which depends on the parameters only */
    __shared__ uint32_t s_doff[PI_MAXSP];
    __shared__ float s_par[PI_MAXSP * PI_NPARAMS];
    __shared__ float s_ogt[PI_MAXSP];
    for (int i = threadIdx.x; i < a.nsp; i += blockDim.x) s_doff[i] = st->doff[i];
    for (int i = threadIdx.x; i < a.nsp * PI_NPARAMS; i += blockDim.x) s_par[i] = a.params[i];
    __syncthreads();
    for (int i = threadIdx.x; i < a.nsp; i += blockDim.x)
        s_ogt[i] = CARBON ? ogt_carbon(s_par + i * PI_NPARAMS) : MM ? ogt_mm(s_par + i * PI_NPARAMS)
                 : IRON ? ogt_iron(s_par + i * PI_NPARAMS) : ogt_of(s_par + i * PI_NPARAMS);
    __syncthreads();
    const uint32_t hhh = (uint32_t)gidx(g, g.Hx, g.Hy, g.Hz);
    const uint32_t stride = K.n * blockDim.x;
    /* the entry and its source are loaded two iterations ahead and its cell
     * (newc of the source) one ahead, so the state and tracer gathers of an
     * entry are issued together at the top of its iteration and no load waits
     * on another within an iteration (one memory latency per entry instead of
     * three). Staging the next entry's state in LDS instead gains nothing: the
     * wait for it also waits for the previous entry's stores. */
    uint32_t e = K.b * blockDim.x + threadIdx.x;
    uint32_t key_n = 0, src_n = 0, c_n = 0, key_nn = 0, src_nn = 0;
    if (e < n) { key_n = a.pkey[e]; src_n = a.psrc[e]; c_n = a.newc[src_n]; }
    if (e < n && n - e > stride) { key_nn = a.pkey[e + stride]; src_nn = a.psrc[e + stride]; }
    for (; e < n; e += stride) {
        const uint32_t key = key_n, src = src_n, c = c_n;
        key_n = key_nn; src_n = src_nn;
        if (n - e > stride) c_n = a.newc[src_n];
        if (n - e > 2 * stride) { key_nn = a.pkey[e + 2 * stride]; src_nn = a.psrc[e + 2 * stride]; }
        const int sp = (int)(key >> PI_SLOT_BITS);
        const uint32_t slot = key & PI_SLOT_MASK;
        const float* q = s_par + sp * PI_NPARAMS;
        const int dvid_type = MM ? 1 : (int)q[CARBON ? (int)PI_C_dvid_type : IRON ? (int)PI_I_dvid_type
                                                                                     : (int)PI_P_dvid_type];
        const float OGT = s_ogt[sp];
        const uint8_t fl = IN.flags[src];
        PState s;
        s.x = IN.x[src]; s.y = IN.y[src]; s.z = IN.z[src];
        s.Sz = IN.Sz[src]; s.Bm = IN.Bm[src]; s.Cq = IN.Cq[src]; s.Nq = IN.Nq[src];
        s.Pq = IN.Pq[src]; s.Chl = IN.Chl[src]; s.gen = IN.gen[src]; s.age = IN.age[src];
        for (int k = 0; k < NEX; ++k) s.ex[k] = IN.ex[k][src];
        /* find_NPT_kernel! (Individuals/utils.jl:14-30), ac = true at gather time */
        Trs t;
        if (CARBON) {                        /* CarbonMode growth reads only par, T and pop */
            t.NH4 = t.NO3 = t.PO4 = t.DOC = t.DFe = 0.0f;
        } else {
            t.NH4 = dmaxf(0.0f, TR_NH4[c]);
            t.NO3 = dmaxf(0.0f, TR_NO3[c]);
            t.PO4 = dmaxf(0.0f, TR_PO4[c]);
            t.DOC = dmaxf(0.0f, TR_DOC[c]);
            t.DFe = dmaxf(0.0f, TR_DFe[c]);
        }
        t.par = PAR[c];
        t.dpar = PAR[c] - PAR0[c];
        t.T = a.temp[c];
        /* shared_graz != 1: pop of this species only (time_step.jl:98-107) */
        t.pop = a.sp_pop ? (float)species_count(a.pkey, a.seg[c], a.seg[c + 1], (uint32_t)sp) : a.pop[c];
        /* grazing! then mortality! (QuotaMode/division_death.jl:2-25, consume_loss.jl:17-33) */
        bool ac = true;
        uint8_t tag = 0;
        float L0 = 0.f, L1 = 0.f, L2 = 0.f, L3 = 0.f, L4 = 0.f, L5 = 0.f, L6 = 0.f, L7 = 0.f, L8 = 0.f, L9 = 0.f;
        if (CARBON && (fl & (F_GRAZ | F_MORT))) {
            /* calc_loss! of CarbonMode (CarbonMode/consume_loss.jl:11-21): Bm to DOC and POC */
            const bool gz = (fl & F_GRAZ) != 0;
            const float fC = gz ? q[PI_C_grazFracC] : q[PI_C_mortFracC];
            L0 = s.Bm * fC * 1.0f;
            L3 = s.Bm * (1.0f - fC) * 1.0f;
            tag = gz ? T_GRAZ : T_MORT;
            ac = false;
        } else if (IRON && (fl & (F_GRAZ | F_MORT))) {
            /* calc_loss! of IronEnergyMode (IronEnergyMode/consume_loss.jl:18-33): the NH4
             * and NO3 reserves as they are, C, N, P and Fe split by the fractions */
            const bool gz = (fl & F_GRAZ) != 0;
            const float fC = gz ? q[PI_I_grazFracC] : q[PI_I_mortFracC];
            const float fN = gz ? q[PI_I_grazFracN] : q[PI_I_mortFracN];
            const float fP = gz ? q[PI_I_grazFracP] : q[PI_I_mortFracP];
            const float fFe = gz ? q[PI_I_grazFracFe] : q[PI_I_mortFracFe];
            const float Fe = s.ex[1] + s.ex[2] + s.ex[4] + s.ex[3];     /* qFe + qFePS + qFeNF + qFeNR */
            L6 = s.Nq * 1.0f;
            L7 = s.ex[0] * 1.0f;
            L0 = (s.Bm + s.Cq) * fC * 1.0f;
            L1 = s.Bm * q[PI_I_R_NC] * fN * 1.0f;
            L2 = (s.Bm * q[PI_I_R_PC] + s.Pq) * fP * 1.0f;
            L8 = Fe * fFe * 1.0f;
            L3 = (s.Bm + s.Cq) * (1.0f - fC) * 1.0f;
            L4 = s.Bm * q[PI_I_R_NC] * (1.0f - fN) * 1.0f;
            L5 = (s.Bm * q[PI_I_R_PC] + s.Pq) * (1.0f - fP) * 1.0f;
            L9 = Fe * (1.0f - fFe) * 1.0f;
            tag = gz ? T_GRAZ : T_MORT;
            ac = false;
        } else if (MM && (fl & (F_GRAZ | F_MORT))) {
            /* calc_loss! of MacroMolecularMode (MacroMolecularMode/consume_loss.jl:17-26):
             * the C, N and P of all the individual's pools */
            const bool gz = (fl & F_GRAZ) != 0;
            const float fC = gz ? q[PI_M_grazFracC] : q[PI_M_mortFracC];
            const float fN = gz ? q[PI_M_grazFracN] : q[PI_M_mortFracN];
            const float fP = gz ? q[PI_M_grazFracP] : q[PI_M_mortFracP];
            const float tC = mm_totC(s), tN = mm_totN(s, q), tP = mm_totP(s, q);
            L0 = tC * fC * 1.0f;
            L1 = tN * fN * 1.0f;
            L2 = tP * fP * 1.0f;
            L3 = tC * (1.0f - fC) * 1.0f;
            L4 = tN * (1.0f - fN) * 1.0f;
            L5 = tP * (1.0f - fP) * 1.0f;
            tag = gz ? T_GRAZ : T_MORT;
            ac = false;
        } else if (fl & (F_GRAZ | F_MORT)) {
            const bool gz = (fl & F_GRAZ) != 0;
            const float fC = gz ? q[PI_P_grazFracC] : q[PI_P_mortFracC];
            const float fN = gz ? q[PI_P_grazFracN] : q[PI_P_mortFracN];
            const float fP = gz ? q[PI_P_grazFracP] : q[PI_P_mortFracP];
            L0 = (s.Bm + s.Cq) * fC * 1.0f;
            L1 = (s.Bm * q[PI_P_R_NC] + s.Nq) * fN * 1.0f;
            L2 = (s.Bm * q[PI_P_R_PC] + s.Pq) * fP * 1.0f;
            L3 = (s.Bm + s.Cq) * (1.0f - fC) * 1.0f;
            L4 = (s.Bm * q[PI_P_R_NC] + s.Nq) * (1.0f - fN) * 1.0f;
            L5 = (s.Bm * q[PI_P_R_PC] + s.Pq) * (1.0f - fP) * 1.0f;
            tag = gz ? T_GRAZ : T_MORT;
            ac = false;
        }
        const bool divider = ac && (fl & F_DVID);
        float U0, U1, U2;
        if (divider) {
            /* divide! (QuotaMode/division_death.jl:28-77): the k-th divider in slot order
             * takes the k-th inactive slot in slot order. */
            const uint32_t w = slot >> 5;
            const uint32_t below = (1u << (slot & 31u)) - 1u;
            const uint32_t r = prefG(a, 0, sp, w) + __popc(a.bmD[(uint32_t)sp * a.nw + w] & below);
            uint32_t lo = 0, hi = (uint32_t)a.nw;           /* largest w' with prefF(w') <= r */
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (prefG(a, 1, sp, mid) <= r) lo = mid; else hi = mid;
            }
            const uint32_t fw = ~a.bmA[(uint32_t)sp * a.nw + lo] & valid_mask(a, lo);
            const int bit = select_bit(fw, r - prefG(a, 1, sp, lo));
            if (bit < 0) {   /* cannot happen after the capacity check; never write out of bounds */
                atomicCAS((uint32_t*)&st->err, 0u, (uint32_t)PI_ERR_HIP);
                continue;
            }
            const uint32_t dslot = lo * 32u + (uint32_t)bit;
            /* daughter = copy before halving; divide_to_half on both */
            PState d = s;
            const long long dr = (long long)sp * a.cap + dslot;
            const bool reborn = a.dstamp[dr] == (int32_t)iteration;
            if (IRON)   /* copy_daughter_individuals! of IronEnergyMode leaves ptc to the slot */
                d.ex[7] = reborn ? a.dbd[dr] : a.tab.ex[7][dr];
            if (CARBON) {
                /* copy_daughter_individuals! of CarbonMode copies x, y, z, Sz, Bm, Chl,
                 * gen: the daughter keeps the slot's Bd (its occupant's, dying in this
                 * step, or the last one's) */
                d.Cq = reborn ? a.dbd[dr] : a.tab.Cq[dr];
                d.Nq = d.Pq = 0.0f;
            }
            const float h = (2.0f - 1.0f) / 2.0f;
            s.Sz = s.Sz * h; s.Bm = s.Bm * h; s.Cq = s.Cq * h; s.Nq = s.Nq * h; s.Pq = s.Pq * h; s.Chl = s.Chl * h;
            s.gen = s.gen + 1.0f; s.age = s.age * (1.0f - 1.0f);
            d.Sz = d.Sz * h; d.Bm = d.Bm * h; d.Cq = d.Cq * h; d.Nq = d.Nq * h; d.Pq = d.Pq * h; d.Chl = d.Chl * h;
            for (int k = 0; k < NEX; ++k)
                if ((HALF >> k) & 1u) { s.ex[k] = s.ex[k] * h; d.ex[k] = d.ex[k] * h; }
            d.gen = d.gen + 1.0f;
            d.age = a.tab.age[dr] * (1.0f - 1.0f);
            /* The daughter's slot kept what find_inds!/find_NPT! wrote into it this
             * step: if its previous occupant was still active then (it dies in this
             * step), the tracers gathered at that occupant's cell and that cell;
             * otherwise zeros and the halo cell (Hx,Hy,Hz). */
            const uint32_t dc = reborn ? a.dcell[dr] : hhh;
            Trs tz;
            if (reborn) {
                tz.NH4 = dmaxf(0.0f, TR_NH4[dc]);
                tz.NO3 = dmaxf(0.0f, TR_NO3[dc]);
                tz.PO4 = dmaxf(0.0f, TR_PO4[dc]);
                tz.DOC = dmaxf(0.0f, TR_DOC[dc]);
                tz.DFe = dmaxf(0.0f, TR_DFe[dc]);
                tz.par = PAR[dc];
                tz.dpar = PAR[dc] - PAR0[dc];
                tz.T = a.temp[dc];
                tz.pop = a.sp_pop ? (float)species_count(a.pkey, a.seg[dc], a.seg[dc + 1], (uint32_t)sp) : a.pop[dc];
            } else {
                tz.NH4 = tz.NO3 = tz.PO4 = tz.DOC = tz.DFe = tz.par = tz.dpar = tz.T = tz.pop = 0.0f;
            }
            pi_uniform3(a.seed, iteration, sp, PI_KIND_DECISION, dslot, &U0, &U1, &U2);
            Rates rd;
            bool gz, mt, dv;
            if (CARBON) growth_carbon(d, tz, q, OGT, true, dt, tmodel, dvid_type, U0, U1, U2, rd, gz, mt, dv);
            else if (MM) growth_mm(d, tz, q, OGT, true, dt, U0, U1, U2, rd, gz, mt, dv);
            else if (IRON) growth_iron(d, tz, q, OGT, true, dt, tmodel, dvid_type, U0, U1, U2, rd, gz, mt, dv);
            else growth(d, tz, q, OGT, true, dt, tmodel, dvid_type, U0, U1, U2, rd, gz, mt, dv);
            if (reborn) {   /* its calc_consume! lands in that cell, at the slot's place */
                a.rcons[dr * PI_RCONS + 0] = (rd.resp - rd.PS) * dt;
                a.rcons[dr * PI_RCONS + 1] = (CARBON || IRON) ? 0.0f : (rd.exu - rd.VDOC) * dt;
                a.rcons[dr * PI_RCONS + 2] = CARBON ? 0.0f : -rd.VNH4 * dt;
                a.rcons[dr * PI_RCONS + 3] = CARBON ? 0.0f : -rd.VNO3 * dt;
                a.rcons[dr * PI_RCONS + 4] = CARBON ? 0.0f : -rd.VPO4 * dt;
                a.rcons[dr * PI_RCONS + 5] = IRON ? -rd.VDOC * dt : 0.0f;   /* DFe (IronEnergyMode: VFe) */
                a.rstamp[dr] = (int32_t)iteration;
            }
            const uint32_t o = n + s_doff[sp] + r;
            if (reborn && a.rdidx) a.rdidx[dr] = o;
            OUT.key[o] = ((uint32_t)sp << PI_SLOT_BITS) | dslot;
            OUT.x[o] = d.x; OUT.y[o] = d.y; OUT.z[o] = d.z;
            OUT.Sz[o] = d.Sz; OUT.Bm[o] = d.Bm; OUT.Cq[o] = d.Cq; OUT.Nq[o] = d.Nq;
            OUT.Pq[o] = d.Pq; OUT.Chl[o] = d.Chl; OUT.gen[o] = d.gen; OUT.age[o] = d.age;
            for (int k = 0; k < NEX; ++k) OUT.ex[k][o] = d.ex[k];
            const uint8_t dfl = (uint8_t)(F_AC | (gz ? F_GRAZ : 0) | (mt ? F_MORT : 0) | (dv ? F_DVID : 0));
            OUT.flags[o] = dfl;
            OUT.cell[o] = dc;
            a.sflag[dr] = dfl;
            if (!reborn) atomicAdd(&a.diag_num[(long long)sp * 4 * g.ncell + hhh], 1.0f);
        }
        pi_uniform3(a.seed, iteration, sp, PI_KIND_DECISION, slot, &U0, &U1, &U2);
        Rates r;
        bool gz, mt, dv;
        if (CARBON) growth_carbon(s, t, q, OGT, ac, dt, tmodel, dvid_type, U0, U1, U2, r, gz, mt, dv);
        else if (MM) growth_mm(s, t, q, OGT, ac, dt, U0, U1, U2, r, gz, mt, dv);
        else if (IRON) growth_iron(s, t, q, OGT, ac, dt, tmodel, dvid_type, U0, U1, U2, r, gz, mt, dv);
        else growth(s, t, q, OGT, ac, dt, tmodel, dvid_type, U0, U1, U2, r, gz, mt, dv);
        if (ac) {
            /* calc_consume! (QuotaMode/consume_loss.jl:2-14, MacroMolecularMode/consume_loss.jl:2-9;
             * CarbonMode: DIC only) */
            tag = T_ALIVE | (divider ? T_DIV : 0);
            L0 = (r.resp - r.PS) * dt;
            L1 = (CARBON || IRON) ? 0.0f : (r.exu - r.VDOC) * dt;
            L2 = CARBON ? 0.0f : -r.VNH4 * dt;
            L3 = CARBON ? 0.0f : -r.VNO3 * dt;
            L4 = CARBON ? 0.0f : -r.VPO4 * dt;
            L5 = IRON ? -r.VDOC * dt : 0.0f;     /* DFe (IronEnergyMode: VFe) */
        } else {
            /* inactive from now on: record its last state in the slot table */
            const long long ts = (long long)sp * a.cap + slot;
            a.tab.x[ts] = s.x; a.tab.y[ts] = s.y; a.tab.z[ts] = s.z;
            a.tab.Sz[ts] = s.Sz; a.tab.Bm[ts] = s.Bm; a.tab.Cq[ts] = s.Cq; a.tab.Nq[ts] = s.Nq;
            a.tab.Pq[ts] = s.Pq; a.tab.Chl[ts] = s.Chl; a.tab.gen[ts] = s.gen; a.tab.age[ts] = s.age;
            for (int k = 0; k < NEX; ++k) a.tab.ex[k][ts] = s.ex[k];
            a.tab.flags[ts] = 0;
            a.tab.cell[ts] = c;
            a.tab.death[ts] = (int32_t)iteration;
        }
        OUT.key[e] = key;
        OUT.x[e] = s.x; OUT.y[e] = s.y; OUT.z[e] = s.z;
        OUT.Sz[e] = s.Sz; OUT.Bm[e] = s.Bm; OUT.Cq[e] = s.Cq; OUT.Nq[e] = s.Nq;
        OUT.Pq[e] = s.Pq; OUT.Chl[e] = s.Chl; OUT.gen[e] = s.gen; OUT.age[e] = s.age;
        for (int k = 0; k < NEX; ++k) OUT.ex[k][e] = s.ex[k];
        const uint8_t ofl = ac ? (uint8_t)(F_AC | (gz ? F_GRAZ : 0) | (mt ? F_MORT : 0) | (dv ? F_DVID : 0)) : (uint8_t)0;
        OUT.flags[e] = ofl;
        OUT.cell[e] = c;
        /* a dying individual leaves its byte as it was (GRAZ or MORT set: not
         * alive for the marks); a daughter may be reborn into that slot.  The
         * byte of a living individual already holds its input flags (every
         * writer of an entry's flags writes its slot's byte too), so it is
         * written only when the decisions changed: one scattered byte store
         * per slot and step cost ~0.14 ms per step at 8.4M individuals (a
         * partial cache-line write-back each), the changes are a few percent */
        if (ac && ofl != (uint8_t)(fl & (F_AC | F_GRAZ | F_MORT | F_DVID)))
            a.sflag[(long long)sp * a.cap + slot] = ofl;
        a.con.v[0][e] = L0; a.con.v[1][e] = L1; a.con.v[2][e] = L2;
        a.con.v[3][e] = L3; a.con.v[4][e] = L4; a.con.v[5][e] = L5;
        if (IRON) { a.con.v[6][e] = L6; a.con.v[7][e] = L7; a.con.v[8][e] = L8; a.con.v[9][e] = L9; }
        a.con.tag[e] = tag;
    }
}
__device__ __forceinline__ void k_physq_body(const PhysArgs& a, Blk K) { k_phys_body<PI_MODE_QUOTA>(a, K); }
__device__ __forceinline__ void k_physc_body(const PhysArgs& a, Blk K) { k_phys_body<PI_MODE_CARBON>(a, K); }
__device__ __forceinline__ void k_physm_body(const PhysArgs& a, Blk K) { k_phys_body<PI_MODE_MACROMOLECULAR>(a, K); }
__device__ __forceinline__ void k_physi_body(const PhysArgs& a, Blk K) { k_phys_body<PI_MODE_IRONENERGY>(a, K); }
PI_KERNEL3(k_phys, k_physq_body, PhysArgs, 256, xcd_block())
PI_KERNEL3(k_physc, k_physc_body, PhysArgs, 256, xcd_block())
PI_KERNEL3(k_physm, k_physm_body, PhysArgs, 256, xcd_block())
PI_KERNEL3(k_physi, k_physi_body, PhysArgs, 256, xcd_block())

/* MacroMolecularMode and IronEnergyMode: the slots inactive through the whole
 * step.  Their plankton_growth! runs with ac = false and the zeros that
 * find_NPT! gathered for them, and changes them: in MacroMolecularMode protein,
 * DNA and RNA synthesis (calc_BS_kernel!), update_quotas_2 and the exudation, in
 * IronEnergyMode the iron allocation (iron_alloc with par = dpar = 0), the dark
 * clock tdark and exEn are not masked by ac.  Rows written in this step (an
 * individual dying now, death = this iteration) already hold their step; live
 * individuals' rows are stale copies (the store holds them) and evolve to no
 * effect. */
template <int MODE>
__device__ __forceinline__ void k_idle_body(const IdleArgs& a, Blk K)
{
    constexpr int NEX = mode_nex(MODE);
    const DevState* st = a.st;
    if (st->err) return;
    const int32_t it = (int32_t)(st->iteration + 1);
    const long long total = a.cap * a.nsp;
    const SlotTable& T = a.tab;
    Trs t;
    t.NH4 = t.NO3 = t.PO4 = t.DOC = t.DFe = t.par = t.dpar = t.T = t.pop = 0.0f;
    for (long long r = K.b * (long long)blockDim.x + threadIdx.x; r < total; r += (long long)K.n * blockDim.x) {
        if ((T.flags[r] & F_AC) || T.death[r] >= it) continue;
        const float* q = a.params + (r / a.cap) * PI_NPARAMS;
        PState s;
        s.x = s.y = s.z = s.gen = 0.0f;
        s.Sz = T.Sz[r]; s.Bm = T.Bm[r]; s.Cq = T.Cq[r]; s.Nq = T.Nq[r]; s.Pq = T.Pq[r];
        s.Chl = T.Chl[r]; s.age = T.age[r];
        for (int k = 0; k < NEX; ++k) s.ex[k] = T.ex[k][r];
        Rates rr;
        bool gz, mt, dv;
        if (MODE == PI_MODE_MACROMOLECULAR) growth_mm(s, t, q, ogt_mm(q), false, a.dt, 1.0f, 1.0f, 1.0f, rr, gz, mt, dv);
        else growth_iron(s, t, q, ogt_iron(q), false, a.dt, 0.0f, 1, 1.0f, 1.0f, 1.0f, rr, gz, mt, dv);
        T.Sz[r] = s.Sz; T.Bm[r] = s.Bm; T.Cq[r] = s.Cq; T.Nq[r] = s.Nq; T.Pq[r] = s.Pq;
        T.Chl[r] = s.Chl; T.age[r] = s.age;
        for (int k = 0; k < NEX; ++k) T.ex[k][r] = s.ex[k];
    }
}
__device__ __forceinline__ void k_idlem_body(const IdleArgs& a, Blk K) { k_idle_body<PI_MODE_MACROMOLECULAR>(a, K); }
__device__ __forceinline__ void k_idlei_body(const IdleArgs& a, Blk K) { k_idle_body<PI_MODE_IRONENERGY>(a, K); }
PI_KERNEL(k_idlem, IdleArgs, 256)
PI_KERNEL(k_idlei, IdleArgs, 256)

/* Ordered per-cell sums: calc_loss! (graz), calc_loss! (mort), calc_consume! per
 * species in slot order (plankton_update.jl:5-29), and the per-species
 * diagnostics num/graz/mort/dvid (diagnostics_kernels.jl:4-12). */

/* the contributions of one entry x of a cell to the ordered sums, per pass:
 * 0 = grazing loss, 1 = mortality loss, 2 = consumption (alive, or a daughter
 * reborn into the slot of this dying entry) */
struct PlkSums {
    float DIC, NH4, NO3, PO4, DOC, DON, DOP, POC, PON, POP, DFe, PFe;
    uint32_t ng, nm, nd, na;
};
/* a loss: v0..v5 to DOC DON DOP POC PON POP; IronEnergyMode also v6..v9 to NH4
 * NO3 DFe PFe_bio (IronEnergyMode/consume_loss.jl:18-33) */
template <bool IRON>
__device__ __forceinline__ void plk_loss(PlkSums& S, const float* v)
{
    S.DOC = S.DOC + v[0]; S.DON = S.DON + v[1]; S.DOP = S.DOP + v[2];
    S.POC = S.POC + v[3]; S.PON = S.PON + v[4]; S.POP = S.POP + v[5];
    if (IRON) { S.NH4 = S.NH4 + v[6]; S.NO3 = S.NO3 + v[7]; S.DFe = S.DFe + v[8]; S.PFe = S.PFe + v[9]; }
}
/* a consumption: DIC DOC NH4 NO3 PO4 (+ DFe in IronEnergyMode) */
template <bool IRON>
__device__ __forceinline__ void plk_cons(PlkSums& S, float c0, float c1, float c2, float c3, float c4, float c5)
{
    S.DIC = S.DIC + c0; S.DOC = S.DOC + c1; S.NH4 = S.NH4 + c2; S.NO3 = S.NO3 + c3; S.PO4 = S.PO4 + c4;
    if (IRON) S.DFe = S.DFe + c5;
}
/* the per-species diagnostics of a cell as atomic additions without return:
 * no load for a later store to wait on, and the counts are integers, so the
 * sums are exact in any order */
__device__ __forceinline__ void plk_diag(const PlkArgs& a, long long c, uint32_t sp, PlkSums& S)
{
    float* d = a.diag + (long long)sp * 4 * a.g.ncell;
    if (S.na) atomicAdd(&d[0 * a.g.ncell + c], (float)S.na);
    if (S.ng) atomicAdd(&d[1 * a.g.ncell + c], (float)S.ng);
    if (S.nm) atomicAdd(&d[2 * a.g.ncell + c], (float)S.nm);
    if (S.nd) atomicAdd(&d[3 * a.g.ncell + c], (float)S.nd);
    S.ng = S.nm = S.nd = S.na = 0;
}

#define PLK_FAST 6   /* segments up to this length are loaded into registers at once */

template <bool IRON>
__device__ __forceinline__ void k_plk_body(const PlkArgs& a, Blk K)
{
    constexpr int NV = IRON ? 10 : 6;        /* loss values per entry */
    if (a.st->err) return;
    const int32_t it = (int32_t)(a.st->iteration + 1);
    const DevGrid& g = a.g;
    const long long stride = (long long)K.n * blockDim.x;
    long long c = K.b * (long long)blockDim.x + threadIdx.x;
    /* the segment bounds are loaded one cell ahead */
    uint32_t s0n = 0, s1n = 0;
    if (c < g.ncell) { s0n = a.seg[c]; s1n = a.seg[c + 1]; }
    for (; c < g.ncell; c += stride) {
        const uint32_t s0 = s0n, s1 = s1n, n = s1 - s0;
        if (c + stride < g.ncell) { s0n = a.seg[c + stride]; s1n = a.seg[c + stride + 1]; }
        a.segp[c] = s0;                              /* the next step's previous segments */
        if (c == g.ncell - 1) a.segp[g.ncell] = s1;
        if (n > LONG_SEG) continue;                  /* k_plk_long */
        PlkSums S{};
        if (n <= PLK_FAST) {
            /* one batch of independent loads, then the three ordered passes per
             * species run on registers (predicated, in slot order) */
            uint32_t key[PLK_FAST];
            uint8_t tag[PLK_FAST];
            float v[PLK_FAST][NV];
#pragma unroll
            for (int x = 0; x < PLK_FAST; ++x) {
                const bool in = (uint32_t)x < n;
                key[x] = in ? a.pkey[s0 + x] : 0xFFFFFFFFu;
                tag[x] = in ? a.con.tag[s0 + x] : (uint8_t)0;
#pragma unroll
                for (int j = 0; j < NV; ++j) v[x][j] = in ? a.con.v[j][s0 + x] : 0.0f;
            }
            uint32_t e = 0;
            while (e < n) {
                uint32_t sp = 0;
#pragma unroll
                for (int x = 0; x < PLK_FAST; ++x) if ((uint32_t)x == e) sp = key[x] >> PI_SLOT_BITS;
                uint32_t e1 = n;
#pragma unroll
                for (int x = PLK_FAST - 1; x >= 0; --x)
                    if ((uint32_t)x > e && (uint32_t)x < n && (key[x] >> PI_SLOT_BITS) != sp) e1 = (uint32_t)x;
#pragma unroll
                for (int x = 0; x < PLK_FAST; ++x)
                    if ((uint32_t)x >= e && (uint32_t)x < e1 && (tag[x] & T_GRAZ)) { plk_loss<IRON>(S, v[x]); ++S.ng; }
#pragma unroll
                for (int x = 0; x < PLK_FAST; ++x)
                    if ((uint32_t)x >= e && (uint32_t)x < e1 && (tag[x] & T_MORT)) { plk_loss<IRON>(S, v[x]); ++S.nm; }
#pragma unroll
                for (int x = 0; x < PLK_FAST; ++x) {
                    if (!((uint32_t)x >= e && (uint32_t)x < e1)) continue;
                    if (tag[x] & T_ALIVE) {
                        plk_cons<IRON>(S, v[x][0], v[x][1], v[x][2], v[x][3], v[x][4], v[x][5]);
                        ++S.na;
                        if (tag[x] & T_DIV) ++S.nd;
                    } else {
                        /* died here this step; a daughter born into its slot consumes here */
                        const long long r = (long long)sp * a.cap + (key[x] & PI_SLOT_MASK);
                        if (a.rstamp[r] == it) {
                            const float* rc = a.rcons + r * PI_RCONS;
                            plk_cons<IRON>(S, rc[0], rc[1], rc[2], rc[3], rc[4], rc[5]);
                            ++S.na;
                        }
                    }
                }
                plk_diag(a, c, sp, S);
                e = e1;
            }
        } else {
            uint32_t e = s0;
            while (e < s1) {
                const uint32_t sp = a.pkey[e] >> PI_SLOT_BITS;
                uint32_t e1 = e + 1;
                while (e1 < s1 && (a.pkey[e1] >> PI_SLOT_BITS) == sp) ++e1;
                float v[NV];
                for (uint32_t x = e; x < e1; ++x) {
                    if (a.con.tag[x] & T_GRAZ) {
                        for (int j = 0; j < NV; ++j) v[j] = a.con.v[j][x];
                        plk_loss<IRON>(S, v); ++S.ng;
                    }
                }
                for (uint32_t x = e; x < e1; ++x) {
                    if (a.con.tag[x] & T_MORT) {
                        for (int j = 0; j < NV; ++j) v[j] = a.con.v[j][x];
                        plk_loss<IRON>(S, v); ++S.nm;
                    }
                }
                for (uint32_t x = e; x < e1; ++x) {
                    const uint8_t t = a.con.tag[x];
                    if (t & T_ALIVE) {
                        plk_cons<IRON>(S, a.con.v[0][x], a.con.v[1][x], a.con.v[2][x], a.con.v[3][x], a.con.v[4][x],
                                       IRON ? a.con.v[5][x] : 0.0f);
                        ++S.na;
                        if (t & T_DIV) ++S.nd;
                    } else {
                        const long long r = (long long)sp * a.cap + (a.pkey[x] & PI_SLOT_MASK);
                        if (a.rstamp[r] == it) {
                            const float* rc = a.rcons + r * PI_RCONS;
                            plk_cons<IRON>(S, rc[0], rc[1], rc[2], rc[3], rc[4], rc[5]);
                            ++S.na;
                        }
                    }
                }
                plk_diag(a, c, sp, S);
                e = e1;
            }
        }
        a.plk[PI_DIC][c] = S.DIC; a.plk[PI_NH4][c] = S.NH4; a.plk[PI_NO3][c] = S.NO3; a.plk[PI_PO4][c] = S.PO4;
        a.plk[PI_DOC][c] = S.DOC; a.plk[PI_DON][c] = S.DON; a.plk[PI_DOP][c] = S.DOP;
        a.plk[PI_POC][c] = S.POC; a.plk[PI_PON][c] = S.PON; a.plk[PI_POP][c] = S.POP;
        if (IRON) { a.plk[PI_DFe][c] = S.DFe; a.plk[PI_PFe_bio][c] = S.PFe; }
    }
}
__device__ __forceinline__ void k_plkq_body(const PlkArgs& a, Blk K) { k_plk_body<false>(a, K); }
__device__ __forceinline__ void k_plki_body(const PlkArgs& a, Blk K) { k_plk_body<true>(a, K); }
PI_KERNEL3(k_plk, k_plkq_body, PlkArgs, 256, xcd_block())
PI_KERNEL3(k_plki, k_plki_body, PlkArgs, 256, xcd_block())


/* The ordered sums of k_plk for the listed long cells, one block per cell.
 * The sums live in lanes 0-9 of wave 0, one tracer per lane in the order
 * DIC NH4 NO3 PO4 DOC DON DOP POC PON POP.  Per species range (slot order):
 * the grazed and the dead entries are listed in order (block prefix sums),
 * then the three passes of the reference run as chains over values staged
 * in LDS, waves 1-3 staging the next chunk while wave 0 adds up the current
 * one.  An entry that contributes nothing to a tracer is staged as +0.0: a
 * sum that starts at +0.0 never becomes −0.0, so adding +0.0 leaves it
 * unchanged, as skipping the entry does. */
template <int U, int NV>
__device__ __forceinline__ void stage_loss(const PlkArgs& a, const uint32_t* L, uint32_t from, uint32_t nL,
                                           uint32_t t0, uint32_t step, float (*dst)[LPAD])
{
    uint32_t x[U];
    float v[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t i = t0 + u * step;
        x[u] = i < LCH && from + i < nL ? L[from + i] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int r = 0; r < NV; ++r) v[u][r] = x[u] != 0xFFFFFFFFu ? a.con.v[r][x[u]] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (x[u] != 0xFFFFFFFFu) {
#pragma unroll
            for (int r = 0; r < NV; ++r) dst[r][t0 + u * step] = v[u][r];
        }
}
/* returns (consuming entries) | (dividing entries) << 16 of the staged ones */
template <int U, int NC>
__device__ __forceinline__ uint32_t stage_cons(const PlkArgs& a, uint32_t sp, int32_t it, uint32_t from, uint32_t e1,
                                               uint32_t t0, uint32_t step, float (*dst)[LPAD])
{
    uint32_t na = 0, nd = 0;
    bool ok[U];
    uint8_t tg[U];
    float v[U][NC];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t i = t0 + u * step;
        ok[u] = i < LCH && from + i < e1;
        tg[u] = ok[u] ? a.con.tag[from + i] : (uint8_t)0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int r = 0; r < NC; ++r) v[u][r] = ok[u] ? a.con.v[r][from + t0 + u * step] : 0.0f;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!ok[u]) continue;
        const uint32_t x = from + t0 + u * step;
        if (tg[u] & T_ALIVE) {
            ++na;
            nd += (tg[u] & T_DIV) ? 1u : 0u;
        } else {
            /* died here this step; a daughter born into its slot consumes here */
            const long long r = (long long)sp * a.cap + (a.pkey[x] & PI_SLOT_MASK);
            const bool rb = a.rstamp[r] == it;
#pragma unroll
            for (int q = 0; q < NC; ++q) v[u][q] = rb ? a.rcons[r * PI_RCONS + q] : 0.0f;
            na += rb ? 1u : 0u;
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) dst[q][t0 + u * step] = v[u][q];
    }
    return na | (nd << 16);
}

template <bool IRON>
__device__ __forceinline__ void k_plk_long_body(const PlkArgs& a, Blk K)
{
    constexpr int NV = IRON ? 10 : 6, NC = IRON ? 6 : 5;   /* staged loss / consumption rows */
    if (a.st->err) return;
    __shared__ __align__(16) float sv[2][NV][LPAD];
    __shared__ uint32_t ws[4];
    __shared__ uint32_t ncnt[2];
    const int32_t it = (int32_t)(a.st->iteration + 1);
    const uint32_t nl = a.st->nlong;
    const uint32_t t = threadIdx.x;
    const long long ncell = a.g.ncell;
    /* lane -> staged row: the loss passes (lanes 4-9: DOC DON DOP POC PON POP
     * <- v0..v5; IronEnergyMode also lanes 1, 2, 10, 11: NH4 NO3 DFe PFe_bio
     * <- v6..v9) and the consumption pass (DIC DOC NH4 NO3 PO4 <- c0..c4;
     * IronEnergyMode also DFe <- c5) */
    const int lrow = t >= 4 && t < 10 ? (int)t - 4
                   : !IRON ? -1 : t == 1 ? 6 : t == 2 ? 7 : t == 10 ? 8 : t == 11 ? 9 : -1;
    const int crow = t == 0 ? 0 : t == 1 ? 2 : t == 2 ? 3 : t == 3 ? 4 : t == 4 ? 1 : (IRON && t == 10) ? 5 : -1;
    for (uint32_t b = K.b; b < nl; b += K.n) {
        const long long c = a.lng[b];
        const uint32_t s0 = a.seg[c], s1 = a.seg[c + 1];
        float S = 0.0f;
        uint32_t e = s0;
        while (e < s1) {
            const uint32_t sp = a.pkey[e] >> PI_SLOT_BITS;
            const uint32_t kend = (sp + 1u) << PI_SLOT_BITS;
            uint32_t lo = e + 1, hi = s1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a.pkey[mid] < kend) lo = mid + 1; else hi = mid;
            }
            const uint32_t e1 = lo;
            if (t == 0) { ncnt[0] = 0; ncnt[1] = 0; }
            /* grazed / dead entries of the range, in order */
            uint32_t nG = 0, nM = 0;
            for (uint32_t base = e; base < e1; base += 1024) {
                uint32_t fg = 0, fm = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t x = base + 4 * t + j;
                    const uint8_t tg = x < e1 ? a.con.tag[x] : (uint8_t)0;
                    fg |= (uint32_t)((tg & T_GRAZ) != 0) << j;
                    fm |= (uint32_t)((tg & T_MORT) != 0) << j;
                }
                uint32_t tot;
                const uint32_t pre = block_excl((uint32_t)__popc(fg) | ((uint32_t)__popc(fm) << 16), ws, tot);
                uint32_t og = nG + (pre & 0xFFFFu), om = nM + (pre >> 16);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t x = base + 4 * t + j;
                    if ((fg >> j) & 1u) a.lg[e + og++] = x;
                    if ((fm >> j) & 1u) a.lm[e + om++] = x;
                }
                nG += tot & 0xFFFFu;
                nM += tot >> 16;
            }
            __syncthreads();
            /* calc_loss! (grazing), then calc_loss! (mortality) */
            for (int pass = 0; pass < 2; ++pass) {
                const uint32_t* L = (pass == 0 ? a.lg : a.lm) + e;
                const uint32_t nL = pass == 0 ? nG : nM;
                if (nL == 0) continue;
                stage_loss<4, NV>(a, L, 0, nL, t, 256, sv[0]);
                __syncthreads();
                uint32_t k = 0;
                for (uint32_t base = 0; base < nL; base += LCH, ++k) {
                    if (t >= 64) { if (base + LCH < nL) stage_loss<6, NV>(a, L, base + LCH, nL, t - 64, 192, sv[(k + 1) & 1]); }
                    else if (lrow >= 0) S = chain_sum(S, sv[k & 1][lrow], min((uint32_t)LCH, nL - base));
                    __syncthreads();
                }
            }
            /* calc_consume! over every entry of the range */
            uint32_t na = 0, nd = 0;
            uint32_t pk = stage_cons<4, NC>(a, sp, it, e, e1, t, 256, sv[0]);
            na += pk & 0xFFFFu; nd += pk >> 16;
            __syncthreads();
            uint32_t k = 0;
            for (uint32_t base = e; base < e1; base += LCH, ++k) {
                if (t >= 64) {
                    if (base + LCH < e1) {
                        pk = stage_cons<6, NC>(a, sp, it, base + LCH, e1, t - 64, 192, sv[(k + 1) & 1]);
                        na += pk & 0xFFFFu; nd += pk >> 16;
                    }
                } else if (crow >= 0) {
                    S = chain_sum(S, sv[k & 1][crow], min((uint32_t)LCH, e1 - base));
                }
                __syncthreads();
            }
            if (na) atomicAdd(&ncnt[0], na);
            if (nd) atomicAdd(&ncnt[1], nd);
            __syncthreads();
            if (t == 0) {   /* plk_diag */
                float* d = a.diag + (long long)sp * 4 * ncell;
                const uint32_t NA = ncnt[0], ND = ncnt[1];
                if (NA) d[0 * ncell + c] = d[0 * ncell + c] + (float)NA;
                if (nG) d[1 * ncell + c] = d[1 * ncell + c] + (float)nG;
                if (nM) d[2 * ncell + c] = d[2 * ncell + c] + (float)nM;
                if (ND) d[3 * ncell + c] = d[3 * ncell + c] + (float)ND;
            }
            __syncthreads();
            e = e1;
        }
        constexpr int QOF[12] = {PI_DIC, PI_NH4, PI_NO3, PI_PO4, PI_DOC, PI_DON, PI_DOP, PI_POC, PI_PON, PI_POP,
                                 PI_DFe, PI_PFe_bio};
#pragma unroll
        for (int l = 0; l < (IRON ? 12 : 10); ++l)
            if (t == (uint32_t)l) a.plk[QOF[l]][c] = S;
    }
}
__device__ __forceinline__ void k_plk_longq_body(const PlkArgs& a, Blk K) { k_plk_long_body<false>(a, K); }
__device__ __forceinline__ void k_plk_longi_body(const PlkArgs& a, Blk K) { k_plk_long_body<true>(a, K); }
PI_KERNEL3(k_plk_long, k_plk_longq_body, PlkArgs, 256, blockIdx.x)
PI_KERNEL3(k_plk_longi, k_plk_longi_body, PlkArgs, 256, blockIdx.x)

/* ------------------------------------------------------ abiotic particles
 * Individuals/Abiotic/particle_interaction.jl:158-267 (particles_from_bcs!)
 * and the abiotic part of TimeStep! (Model/time_step.jl:21-38, 152-159). */

/* the area of face f of cell (i, j, k) (1-based padded): Az top/bottom, Ax
 * east/west, Ay north/south (Grids/areas_volumes_spacings.jl) */
__device__ __forceinline__ float face_area(const DevGrid& g, int f, int i, int j, int k, long long c)
{
    if (f == PI_BC_TOP || f == PI_BC_BOTTOM) return g.gm ? g.mAz[col2(g, i, j)] : g.dx * g.dy;
    if (f == PI_BC_EAST || f == PI_BC_WEST) return g.gm ? g.mAx[c] : g.dy * dzF_at(g, k);
    return g.gm ? g.mAy[c] : g.dx * dzF_at(g, k);
}

/* calc_particle_bcs! (:158-236): the cells next to each face with a
 * condition get isless(u·sz_min·Nsuper, bc·ΔT·A); the faces are evaluated in
 * the order top, bottom, east, west, north, south, so at an edge or a corner
 * the last of them decides.  u is the cell's draw of rand!(rnd_3d). */
__device__ __forceinline__ void k_abio_mark_body(const AbioArgs& a, Blk K)
{
    const DevState* st = a.adv.st;
    if (st->err) return;
    const DevGrid& g = a.adv.g;
    const long long n = (long long)a.nsa * a.ncell;
    const int it = (int)(st->iteration + 1);
    const uint32_t sxy = (uint32_t)g.Px * (uint32_t)g.Py;
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)K.n * blockDim.x) {
        const int sa = (int)(t / a.ncell);
        const long long c = t - (long long)sa * a.ncell;
        const uint32_t m = (a.bcm >> (6 * sa)) & 63u;
        uint32_t flag = 0;
        if (m) {
            const int i = (int)(c % g.Px) + 1, j = (int)((c / g.Px) % g.Py) + 1, k = (int)(c / sxy) + 1;
            const int ii = i - g.Hx, jj = j - g.Hy, kk = k - g.Hz;          /* interior, 1-based */
            if (ii >= 1 && ii <= g.Nx && jj >= 1 && jj <= g.Ny && kk >= 1 && kk <= g.Nz) {
                static constexpr int order[6] = {PI_BC_TOP, PI_BC_BOTTOM, PI_BC_EAST, PI_BC_WEST, PI_BC_NORTH, PI_BC_SOUTH};
                float u, u1, u2;
                pi_uniform3(a.adv.seed, it, PI_ABIO_SPECIES + sa, PI_KIND_ABIO_BC, c, &u, &u1, &u2);
                const float lhs = (u * a.p[sa][PI_A_sz_min]) * a.p[sa][PI_A_Nsuper];
#pragma unroll
                for (int o = 0; o < 6; ++o) {
                    const int f = order[o];
                    if (!((m >> f) & 1u)) continue;
                    const bool on = f == PI_BC_TOP ? kk == 1 : f == PI_BC_BOTTOM ? kk == g.Nz : f == PI_BC_EAST ? ii == g.Nx
                                  : f == PI_BC_WEST ? ii == 1 : f == PI_BC_NORTH ? jj == g.Ny : jj == 1;
                    if (!on) continue;
                    /* getbc at the face position: (i, j) top/bottom, (j, k) east/west, (i, k) north/south */
                    const int p0 = f <= PI_BC_EAST ? jj : ii, p1 = f <= PI_BC_NORTH ? kk : jj;
                    const int n0 = f <= PI_BC_EAST ? g.Ny : g.Nx, n1 = f <= PI_BC_NORTH ? g.Nz : g.Ny;
                    const int kind = a.bc->kind[sa][f];
                    float b;
                    if (kind == PI_BCK_CONST) b = a.bc->val[sa][f];
                    else {
                        const long long tt = kind == PI_BCK_SERIES ? (long long)(it - 1) : 0ll;
                        b = a.bc->arr[sa][f][(long long)(p0 - 1) + (long long)n0 * ((long long)(p1 - 1) + (long long)n1 * tt)];
                    }
                    flag = jl_islessf(lhs, (b * a.adv.dt) * face_area(g, f, i, j, k, c)) ? 1u : 0u;
                }
            }
        }
        a.cflag[t] = flag;
    }
}
PI_KERNEL(k_abio_mark, AbioArgs, 256)

/* findall(isequal(false), abiotic.ac): the free slots, flagged */
__device__ __forceinline__ void k_abio_free_body(const AbioArgs& a, Blk K)
{
    if (a.adv.st->err) return;
    const long long n = (long long)a.nsa * a.cap;
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)K.n * blockDim.x)
        a.fflag[t] = a.ac[t] ? 0u : 1u;
}
PI_KERNEL(k_abio_free, AbioArgs, 256)

/* the free slots of each species in ascending order (deactive_ind) */
__device__ __forceinline__ void k_abio_list_body(const AbioArgs& a, Blk K)
{
    if (a.adv.st->err) return;
    const long long n = (long long)a.nsa * a.cap;
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)K.n * blockDim.x) {
        if (a.ac[t]) continue;
        const long long sa = t / a.cap;
        a.flist[sa * a.cap + (a.fpre[t] - a.fpre[sa * a.cap])] = (uint32_t)(t - sa * a.cap);
    }
}
PI_KERNEL(k_abio_list, AbioArgs, 256)

/* length(inds) > length(deactive_ind) throws (:256-259): checked before any
 * buffer of the step changes */
__device__ __forceinline__ void k_abio_check_body(const AbioArgs& a, Blk K)
{
    DevState* st = a.adv.st;
    if (threadIdx.x != 0 || st->err) return;
    for (int sa = 0; sa < a.nsa; ++sa) {
        const uint32_t make = a.cpre[(long long)(sa + 1) * a.ncell] - a.cpre[(long long)sa * a.ncell];
        const uint32_t free = a.fpre[(long long)(sa + 1) * a.cap] - a.fpre[(long long)sa * a.cap];
        if (make > free) { st->err = PI_ERR_CAPACITY; st->err_sp = PI_ERR_SP_ABIO + (uint32_t)sa; return; }
    }
}
PI_KERNEL(k_abio_check, AbioArgs, 64)

/* copy_abiotic_particle_from_field! (:238-247): the k-th flagged cell in
 * column-major order fills the k-th free slot, at the cell's corner + 0.6 */
__device__ __forceinline__ void k_abio_spawn_body(const AbioArgs& a, Blk K)
{
    if (a.adv.st->err) return;
    const DevGrid& g = a.adv.g;
    const long long n = (long long)a.nsa * a.ncell;
    const uint32_t sxy = (uint32_t)g.Px * (uint32_t)g.Py;
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)K.n * blockDim.x) {
        if (!a.cflag[t]) continue;
        const long long sa = t / a.ncell, c = t - sa * a.ncell;
        const uint32_t kth = a.cpre[t] - a.cpre[sa * a.ncell];
        const long long r = sa * a.cap + a.flist[sa * a.cap + kth];
        const int i = (int)(c % g.Px) + 1, j = (int)((c / g.Px) % g.Py) + 1, k = (int)(c / sxy) + 1;
        a.ac[r] = 1;
        a.x[r] = (float)(i - g.Hx) - 0.4f;
        a.y[r] = (float)(j - g.Hy) - 0.4f;
        a.z[r] = (float)(k - g.Hz) - 0.4f;
    }
}
PI_KERNEL(k_abio_spawn, AbioArgs, 256)

/* particle_advection!, particle_diffusion! and find_inds! of every abiotic
 * slot (time_step.jl:27-36), then diags_proc! of num (:152-159); inactive
 * slots keep ±0 positions and the halo cell (Hx, Hy, Hz) */
template <bool GM>
__device__ __forceinline__ void k_abio_adv_impl(const AbioArgs& a, Blk K)
{
    const DevState* st = a.adv.st;
    if (st->err) return;
    const DevGrid& g = a.adv.g;
    const long long n = (long long)a.nsa * a.cap;
    const long long it = st->iteration + 1;
    for (long long t = K.b * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)K.n * blockDim.x) {
        const long long sa = t / a.cap, slot = t - sa * a.cap;
        if (!a.ac[t]) {
            a.x[t] = copysignf(0.0f, a.x[t]); a.y[t] = copysignf(0.0f, a.y[t]); a.z[t] = copysignf(0.0f, a.z[t]);
            a.xi[t] = g.Hx; a.yi[t] = g.Hy; a.zi[t] = g.Hz;
            continue;
        }
        float r0, r1, r2, x, y, z;
        pi_uniform3(a.adv.seed, it, PI_ABIO_SPECIES + (int)sa, PI_KIND_DIFFUSION, slot, &r0, &r1, &r2);
        move<GM>(a.adv, g, a.x[t], a.y[t], a.z[t], r0, r1, r2, a.xi[t], a.yi[t], a.zi[t], x, y, z);
        int xi = (int)(x + 1.0f) + g.Hx, yi = (int)(y + 1.0f) + g.Hy, zi = (int)(z + 1.0f) + g.Hz;
        a.x[t] = x; a.y[t] = y; a.z[t] = z;
        a.xi[t] = xi; a.yi[t] = yi; a.zi[t] = zi;
        xi = min(max(xi, 1), g.Px); yi = min(max(yi, 1), g.Py); zi = min(max(zi, 1), g.Pz);
        if (!a.num_late) atomicAdd(&a.num[sa * a.ncell + gidx(g, xi, yi, zi)], 1.0f);
    }
}
__device__ __forceinline__ void k_abio_adv_body(const AbioArgs& a, Blk K) { k_abio_adv_impl<false>(a, K); }
__device__ __forceinline__ void k_abio_adv_gm_body(const AbioArgs& a, Blk K) { k_abio_adv_impl<true>(a, K); }
PI_KERNEL(k_abio_adv, AbioArgs, 256)
PI_KERNEL(k_abio_adv_gm, AbioArgs, 256)

/* ------------------------------------ phytoplankton-particle interaction
 * Abiotic/particle_interaction.jl:1-145 and time_step.jl:120-159, run after the
 * physiology of every species on the next step's store (k_phys has written it):
 * the living individuals at their cells, a daughter reborn into a dying
 * individual's slot at that slot's cell, other daughters at the halo cell
 * (Hx, Hy, Hz) where no particle is.  The reference pairs every individual with
 * every particle (calc_interaction_kernel!) and keeps those of the same cell;
 * here the pair's particles are listed per cell (count, scan, fill) and one
 * thread per cell walks its particles in slot order against the species'
 * individuals of the cell, so the work is linear in the slots and cells plus,
 * per cell, particles x individuals of the cell.  The reference's order matters
 * only within a cell (a particle can join an individual of its own cell only),
 * so cells run in parallel.  Releases are ranked by device-wide scans.  Every
 * kernel below handles the pair a.q of a.pint / a.prel. */

__device__ __forceinline__ uint32_t int_acell(const InteractArgs& a, long long r)
{   /* the particle's cell (find_inds! of this step; stale for inactive slots) */
    const DevGrid& g = a.g;
    const int xi = min(max((int)a.xi[r], 1), g.Px), yi = min(max((int)a.yi[r], 1), g.Py), zi = min(max((int)a.zi[r], 1), g.Pz);
    return (uint32_t)gidx(g, xi, yi, zi);
}
__device__ __forceinline__ uint32_t int_nout(const InteractArgs& a, const DevState* st)
{   /* rows of the next store: the cell segments and the daughters after them */
    uint32_t n = a.seg[a.g.ncell];
    for (int s = 0; s < a.nsp; ++s) n += st->dtot[s];
    return n;
}

/* reset_interaction!: the per-cell counts and fill positions */
__device__ __forceinline__ void k_int_zero_body(const InteractArgs& a, Blk K)
{
    if (a.st->err) return;
    for (long long c = K.b * (long long)blockDim.x + threadIdx.x; c < a.g.ncell; c += (long long)K.n * blockDim.x) {
        a.cnt[c] = 0u;
        a.fill[c] = 0u;
    }
}
__device__ __forceinline__ void k_int_count_body(const InteractArgs& a, Blk K)
{
    if (a.st->err) return;
    const long long r0 = (long long)a.pint[a.q][1] * a.cap;
    for (long long k = K.b * (long long)blockDim.x + threadIdx.x; k < a.cap; k += (long long)K.n * blockDim.x)
        if (a.ac[r0 + k]) atomicAdd(&a.cnt[int_acell(a, r0 + k)], 1u);
}
__device__ __forceinline__ void k_int_list_body(const InteractArgs& a, Blk K)
{
    if (a.st->err) return;
    const long long r0 = (long long)a.pint[a.q][1] * a.cap;
    for (long long k = K.b * (long long)blockDim.x + threadIdx.x; k < a.cap; k += (long long)K.n * blockDim.x)
        if (a.ac[r0 + k]) {
            const uint32_t c = int_acell(a, r0 + k);
            a.list[a.off[c] + atomicAdd(&a.fill[c], 1u)] = (uint32_t)k;
        }
}
/* calc_interaction_kernel! and merge_particle_kernel! (:22-72), one cell per thread */
__device__ __forceinline__ void k_int_cells_body(const InteractArgs& a, Blk K)
{
    DevState* st = a.st;
    if (st->err) return;
    const DevGrid& g = a.g;
    const int sp = a.pint[a.q][0], sa = a.pint[a.q][1];
    const long long r0 = (long long)sa * a.cap;
    const float Rd = a.ap[sa][PI_A_Rd];
    const float maxptc = a.params[sp * PI_NPARAMS + PI_I_max_ptc];
    const int32_t it = (int32_t)(st->iteration + 1);
    const GStore O = gstore(st->store[1 - st->scur]);
    const uint32_t sxy = (uint32_t)g.Px * (uint32_t)g.Py;
    for (long long c = K.b * (long long)blockDim.x + threadIdx.x; c < g.ncell; c += (long long)K.n * blockDim.x) {
        const uint32_t m = a.cnt[c];
        if (!m) continue;
        uint32_t* L = a.list + a.off[c];
        for (uint32_t q = 1; q < m; ++q) {          /* the cell's particles in slot order */
            const uint32_t v = L[q];
            uint32_t j = q;
            while (j > 0 && L[j - 1] > v) { L[j] = L[j - 1]; --j; }
            L[j] = v;
        }
        /* this species' individuals of the cell (the segment is ordered by species, slot) */
        uint32_t e0 = a.seg[c], e1 = a.seg[c + 1];
        while (e0 < e1 && (O.key[e0] >> PI_SLOT_BITS) < (uint32_t)sp) ++e0;
        uint32_t e2 = e0;
        while (e2 < e1 && (O.key[e2] >> PI_SLOT_BITS) == (uint32_t)sp) ++e2;
        if (e0 == e2) continue;
        const int ci = (int)(c % g.Px) + 1, cj = (int)((c / g.Px) % g.Py) + 1, ck = (int)(c / sxy) + 1;
        const float DX = g.gm ? g.mdxF[col2(g, ci, cj)] : g.dx;       /* ΔxF, ΔyF, ΔzF of the cell */
        const float DY = g.gm ? g.mdyF[col2(g, ci, cj)] : g.dy;
        const float DZ = dzF_at(g, ck);
        for (uint32_t q = 0; q < m; ++q) {
            const long long r = r0 + L[q];
            uint32_t cidx[PI_MAXCAND];                    /* intac[:, k]: the candidate's row, ~0u = none */
            for (int u = 0; u < a.mc; ++u) cidx[u] = ~0u;
            const float px = a.x[r], py = a.y[r], pz = a.z[r];
            for (uint32_t e = e0; e < e2; ++e) {
                const uint32_t slot = O.key[e] & PI_SLOT_MASK;
                uint32_t o;
                if (O.flags[e] & F_AC) o = e;
                else if (a.rstamp[(long long)sp * a.cap + slot] == it) o = a.rdidx[(long long)sp * a.cap + slot];
                else continue;
                const float dx = (O.x[o] - px) * DX, dy = (O.y[o] - py) * DY, dz = (O.z[o] - pz) * DZ;
                const float dist = sqrtf(pi_pow2f(dx) + pi_pow2f(dy) + pi_pow2f(dz));
                if (dist < Rd) cidx[(slot + 1u) % (uint32_t)a.mc] = o;   /* slot = (i % max_candidates) + 1 */
            }
            for (int u = 0; u < a.mc; ++u) {              /* merge_particle_kernel! */
                const uint32_t o = cidx[u];
                if (o != ~0u && O.ex[7][o] < maxptc) {
                    O.ex[7][o] = O.ex[7][o] + 1.0f;
                    a.ac[r] = 0;
                    break;
                }
            }
        }
    }
}

/* particle_release! (:86-138): flags of the releasers by slot and of the free
 * particle slots, two scans, the capacity check, the free-slot list, then the
 * k-th releaser in slot order fills the k-th free slot; rnd.x is this call's
 * draw, rnd.y and rnd.z still hold the last species' mortality and division
 * draws of get_probability! */
__device__ __forceinline__ void k_rel_zero_body(const InteractArgs& a, Blk K)
{
    if (a.st->err) return;
    const long long r0 = (long long)a.prel[a.q][1] * a.cap;
    for (long long s = K.b * (long long)blockDim.x + threadIdx.x; s < a.cap; s += (long long)K.n * blockDim.x) {
        a.rflag[s] = 0u;
        a.fflag[s] = a.ac[r0 + s] ? 0u : 1u;
    }
}
__device__ __forceinline__ void k_rel_flag_body(const InteractArgs& a, Blk K)
{
    const DevState* st = a.st;
    if (st->err) return;
    const int sp = a.prel[a.q][0];
    const float rP = a.ap[a.prel[a.q][1]][PI_A_release_P];
    const int32_t it = (int32_t)(st->iteration + 1);
    const GStore O = gstore(st->store[1 - st->scur]);
    const uint32_t nout = int_nout(a, st);
    for (long long e = K.b * (long long)blockDim.x + threadIdx.x; e < nout; e += (long long)K.n * blockDim.x) {
        const uint32_t key = O.key[e];
        if ((key >> PI_SLOT_BITS) != (uint32_t)sp || !(O.flags[e] & F_AC)) continue;
        const uint32_t slot = key & PI_SLOT_MASK;
        float u0, u1, u2;
        pi_uniform3(a.seed, it, PI_RELEASE_SPECIES + a.q, PI_KIND_RELEASE, slot, &u0, &u1, &u2);
        if (jl_islessf(u0, rP * a.dt * O.ex[7][e])) a.rflag[slot] = 1u;
    }
}
__device__ __forceinline__ void k_rel_check_body(const InteractArgs& a, Blk K)
{
    (void)K;
    DevState* st = a.st;
    if (threadIdx.x == 0 && !st->err && a.rrank[a.cap] > a.fpre[a.cap]) {
        st->err = PI_ERR_CAPACITY;
        st->err_sp = PI_ERR_SP_RELEASE + (uint32_t)a.prel[a.q][1];
    }
}
__device__ __forceinline__ void k_rel_list_body(const InteractArgs& a, Blk K)
{
    if (a.st->err) return;
    for (long long s = K.b * (long long)blockDim.x + threadIdx.x; s < a.cap; s += (long long)K.n * blockDim.x)
        if (a.fflag[s]) a.flist[a.fpre[s]] = (uint32_t)s;
}
__device__ __forceinline__ void k_rel_write_body(const InteractArgs& a, Blk K)
{
    const DevState* st = a.st;
    if (st->err) return;
    const int sp = a.prel[a.q][0];
    const long long r0 = (long long)a.prel[a.q][1] * a.cap;
    const int32_t it = (int32_t)(st->iteration + 1);
    const GStore O = gstore(st->store[1 - st->scur]);
    const uint32_t nout = int_nout(a, st);
    for (long long e = K.b * (long long)blockDim.x + threadIdx.x; e < nout; e += (long long)K.n * blockDim.x) {
        const uint32_t key = O.key[e];
        if ((key >> PI_SLOT_BITS) != (uint32_t)sp || !(O.flags[e] & F_AC)) continue;
        const uint32_t slot = key & PI_SLOT_MASK;
        if (!a.rflag[slot]) continue;
        float u0, u1, u2, v0, v1, v2;
        pi_uniform3(a.seed, it, PI_RELEASE_SPECIES + a.q, PI_KIND_RELEASE, slot, &u0, &u1, &u2);
        pi_uniform3(a.seed, it, a.nsp - 1, PI_KIND_DECISION, slot, &v0, &v1, &v2);
        const long long r = r0 + a.flist[a.rrank[slot]];
        a.ac[r] = 1;
        a.x[r] = O.x[e] + u0;
        a.y[r] = O.y[e] + v1;
        a.z[r] = O.z[e] + v2;
        O.ex[7][e] = O.ex[7][e] - 1.0f;
    }
}
/* diags_proc! of ptc for every species (time_step.jl:140-149) */
__device__ __forceinline__ void k_int_diag_body(const InteractArgs& a, Blk K)
{
    const DevState* st = a.st;
    if (st->err) return;
    const GStore O = gstore(st->store[1 - st->scur]);
    const uint32_t nout = int_nout(a, st);
    for (long long e = K.b * (long long)blockDim.x + threadIdx.x; e < nout; e += (long long)K.n * blockDim.x) {
        if (!(O.flags[e] & F_AC)) continue;
        const float v = O.ex[7][e];
        if (v != 0.0f) atomicAdd(&a.diag_ptc[(long long)(O.key[e] >> PI_SLOT_BITS) * a.g.ncell + O.cell[e]], v);
    }
}
/* diags_proc! of num for every abiotic species (:152-159), after the pairs */
__device__ __forceinline__ void k_int_num_body(const InteractArgs& a, Blk K)
{
    if (a.st->err) return;
    for (long long r = K.b * (long long)blockDim.x + threadIdx.x; r < (long long)a.nsa * a.cap; r += (long long)K.n * blockDim.x)
        if (a.ac[r]) atomicAdd(&a.num[(r / a.cap) * a.g.ncell + int_acell(a, r)], 1.0f);
}
PI_KERNEL(k_int_zero, InteractArgs, 256)
PI_KERNEL(k_int_count, InteractArgs, 256)
PI_KERNEL(k_int_list, InteractArgs, 256)
PI_KERNEL(k_int_cells, InteractArgs, 256)
PI_KERNEL(k_rel_zero, InteractArgs, 256)
PI_KERNEL(k_rel_flag, InteractArgs, 256)
PI_KERNEL(k_rel_check, InteractArgs, 64)
PI_KERNEL(k_rel_list, InteractArgs, 256)
PI_KERNEL(k_rel_write, InteractArgs, 256)
PI_KERNEL(k_int_diag, InteractArgs, 256)
PI_KERNEL(k_int_num, InteractArgs, 256)

/* ------------------------------------------------------------- registry */
void reg_particles(KInfo* k)
{
    PI_REG(k, K_ABIO_MARK, k_abio_mark, AbioArgs, 256, true);
    PI_REG(k, K_ABIO_FREE, k_abio_free, AbioArgs, 256, true);
    PI_REG(k, K_ABIO_LIST, k_abio_list, AbioArgs, 256, true);
    PI_REG(k, K_ABIO_CHECK, k_abio_check, AbioArgs, 64, false);
    PI_REG(k, K_ABIO_SPAWN, k_abio_spawn, AbioArgs, 256, true);
    PI_REG(k, K_ABIO_ADV, k_abio_adv, AbioArgs, 256, true);
    PI_REG(k, K_ABIO_ADV_GM, k_abio_adv_gm, AbioArgs, 256, true);
    PI_REG(k, K_ADV, k_adv, AdvArgs, 256, true);
    PI_REG(k, K_ADV_GM, k_adv_gm, AdvArgs, 256, true);
    PI_REG(k, K_MARKS, k_marks, MarksArgs, 256, true);
    PI_REG(k, K_CELLCOUNT, k_cellcount, BinArgs, 256, true);
    PI_REG(k, K_PLACE_STAY, k_place_stay, BinArgs, 256, true);
    PI_REG(k, K_PLACE_SMALL, k_place_small, BinArgs, 256, true);
    PI_REG(k, K_MIXSORT, k_mixsort, BinArgs, 256, true);
    PI_REG(k, K_BIGCELLS, k_bigcells, BinArgs, 256, true);
    PI_REG(k, K_CELLSUM, k_cellsum, BinArgs, 256, true);
    PI_REG(k, K_CELLSUM_LONG, k_cellsum_long, BinArgs, 256, true);
    PI_REG(k, K_PAR, k_par, ParArgs, 256, true);
    PI_REG(k, K_PHYS, k_phys, PhysArgs, 256, true);
    PI_REG(k, K_PHYSC, k_physc, PhysArgs, 256, true);
    PI_REG(k, K_PHYSM, k_physm, PhysArgs, 256, true);
    PI_REG(k, K_IDLEM, k_idlem, IdleArgs, 256, true);
    PI_REG(k, K_PHYSI, k_physi, PhysArgs, 256, true);
    PI_REG(k, K_IDLEI, k_idlei, IdleArgs, 256, true);
    PI_REG(k, K_PLK, k_plk, PlkArgs, 256, true);
    PI_REG(k, K_PLK_LONG, k_plk_long, PlkArgs, 256, true);
    PI_REG(k, K_PLKI, k_plki, PlkArgs, 256, true);
    PI_REG(k, K_PLK_LONGI, k_plk_longi, PlkArgs, 256, true);
    PI_REG(k, K_INT_ZERO, k_int_zero, InteractArgs, 256, true);
    PI_REG(k, K_INT_COUNT, k_int_count, InteractArgs, 256, true);
    PI_REG(k, K_INT_LIST, k_int_list, InteractArgs, 256, true);
    PI_REG(k, K_INT_CELLS, k_int_cells, InteractArgs, 256, true);
    PI_REG(k, K_REL_ZERO, k_rel_zero, InteractArgs, 256, true);
    PI_REG(k, K_REL_FLAG, k_rel_flag, InteractArgs, 256, true);
    PI_REG(k, K_REL_CHECK, k_rel_check, InteractArgs, 64, false);
    PI_REG(k, K_REL_LIST, k_rel_list, InteractArgs, 256, true);
    PI_REG(k, K_REL_WRITE, k_rel_write, InteractArgs, 256, true);
    PI_REG(k, K_INT_DIAG, k_int_diag, InteractArgs, 256, true);
    PI_REG(k, K_INT_NUM, k_int_num, InteractArgs, 256, true);
}
