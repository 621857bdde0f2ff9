/*
 * k_tracers.hip - tracer_update! (Biogeochemistry/tracer_update.jl:22-43) on gfx950.
 *
 *   k_sweep_x / k_sweep_y  one DST3FL dimension of tracer_advection!
 *                          (Advection/multi_dim_adv.jl:111-135) for all 14
 *                          tracers per lane; face quantities (CFL, d0, d1,
 *                          transport) are computed once and shared by the 14.
 *   k_fill_line            halo fill of one dimension along lines, as the
 *                          reference's fill_halo_* (HaloRegions/halo_regions.jl)
 *                          with broadcast unaliasing (each line is owned by one
 *                          lane, which reads its sources before writing).
 *   k_final                z sweep + Gcs = temp - c + diffusion + sinking +
 *                          forcing + apply_tendency, writing the next tracers.
 *
 * The face flux at the last interior face is computed from the halo values for
 * Periodic directions (identical to the reference's Gc halo copy of the first
 * face) and is zero for Bounded ones (fill_halo_east_Gc!/..., halo_regions.jl:22-24).
 */
#include "pi_device.h"
#include "k_args.h"

__device__ __forceinline__ float lmask(const DevGrid& g, long long outer, long long inner, float x)
{
    if (!g.has_land) return x;
    return jl_maskf(g.lm[outer] != 0, jl_maskf(g.lm[inner] != 0, x));
}

/* DST3FL.jl:23-57 */
__device__ __forceinline__ float theta(float dn, float dd)
{
    const float tmax = 1.0e20f;
    return fabsf(dd) * tmax <= fabsf(dn) ? copysignf(tmax, dn * dd) : dn / dd;
}
__device__ __forceinline__ float psi(float C, float d0, float d1, float th)
{
    return jl_maxf(0.0f, jl_minf(jl_minf(1.0f, d0 + d1 * th), th * (1.0f - C) / (C + 1.0e-20f)));
}

/* adv_flux_{x,y} at face f (the face between f-1 and f) along a line with stride
 * `st` (1 for x, Px for y); c is the field being swept. DST3FL.jl:60-91 */
__device__ __forceinline__ float flux_xy(const DevGrid& g, const float* __restrict__ c, long long f, long long st,
                                         float C, float d0, float d1, float T)
{
    const float cm2 = c[f - 2 * st], cm1 = c[f - st], c0 = c[f], cp1 = c[f + st];
    const float dd0 = lmask(g, f - st, f, c0 - cm1);
    const float ddm = lmask(g, f - 2 * st, f - st, cm1 - cm2);
    const float ddp = lmask(g, f, f + st, cp1 - c0);
    const float pp = psi(C, d0, d1, theta(ddm, dd0));
    const float pm = psi(C, d0, d1, theta(ddp, dd0));
    return 0.5f * (T + fabsf(T)) * (cm1 + pp * dd0) + 0.5f * (T - fabsf(T)) * (c0 - pm * dd0);
}
/* adv_flux_z at face k (between k-1 above and k below), stride sz. DST3FL.jl:93-95 */
__device__ __forceinline__ float flux_z(const DevGrid& g, const float* __restrict__ c, long long f, long long sz,
                                        float C, float d0, float d1, float T)
{
    const float cm2 = c[f - 2 * sz], cm1 = c[f - sz], c0 = c[f], cp1 = c[f + sz];
    const float dd0 = lmask(g, f - sz, f, cm1 - c0);
    const float ddm = lmask(g, f - 2 * sz, f - sz, cm2 - cm1);
    const float ddp = lmask(g, f, f + sz, c0 - cp1);
    const float pp = psi(C, d0, d1, theta(ddm, dd0));   /* Ψz⁺ */
    const float pm = psi(C, d0, d1, theta(ddp, dd0));   /* Ψz⁻ */
    return 0.5f * (T + fabsf(T)) * (c0 + pm * dd0) + 0.5f * (T - fabsf(T)) * (cm1 - pp * dd0);
}


/* calc_Gcsˣ!/multi_dim_x! and calc_Gcsʸ!/multi_dim_y! for one interior cell */
__global__ void __launch_bounds__(256) k_sweep(SweepArgs a)
{
    const DevState* st = a.st;
    if (st->err) return;
    const DevGrid& g = a.g;
    const long long nint = (long long)g.Nx * g.Ny * g.Nz;
    const float* const* C0 = st->tr[st->tcur];
    const long long stride = a.dim == 0 ? 1 : g.Px;
    const float dt = a.dt;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nint; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.Nx) + 1 + g.Hx;
        const int j = (int)((t / g.Nx) % g.Ny) + 1 + g.Hy;
        const int k = (int)(t / ((long long)g.Nx * g.Ny)) + 1 + g.Hz;
        const long long c = gidx(g, i, j, k);
        const int pos = a.dim == 0 ? i : j;
        const int N = a.dim == 0 ? g.Nx : g.Ny, H = a.dim == 0 ? g.Hx : g.Hy;
        const int topo = a.dim == 0 ? g.tx : g.ty;
        const float D = a.dim == 0 ? g.dx : g.dy;
        const float A = (a.dim == 0 ? g.dy : g.dx) * dzF_at(g, k);            /* Ax / Ay */
        const float vol = g.dx * g.dy * dzF_at(g, k);
        const float u0 = a.vel[c], u1 = a.vel[c + stride];
        const float Cf0 = fabsf(u0 * dt / D), Cf1 = fabsf(u1 * dt / D);
        const float T0 = A * u0, T1 = A * u1;
        const float a0 = (2.0f - Cf0) * (1.0f - Cf0) / 6.0f, b0 = (1.0f - Cf0 * Cf0) / 6.0f;
        const float a1 = (2.0f - Cf1) * (1.0f - Cf1) / 6.0f, b1 = (1.0f - Cf1 * Cf1) / 6.0f;
        const bool last_bounded = (topo == PI_BOUNDED) && (pos == N + H);
        const float dT = lmask(g, c, c + stride, T1 - T0);
        const float coef = dt / vol;
        for (int q = 0; q < PI_NTRACERS; ++q) {
            const float* src = a.prev[q] ? a.prev[q] : C0[q];
            const float F0 = flux_xy(g, src, c, stride, Cf0, a0, b0, T0);
            const float F1 = last_bounded ? 0.0f : flux_xy(g, src, c + stride, stride, Cf1, a1, b1, T1);
            const float dG = lmask(g, c, c + stride, F1 - F0);
            a.out[q][c] = src[c] - coef * (dG - C0[q][c] * dT);
        }
    }
}

/* fill_halo_{west,east} / {south,north} / {top,bottom} of one dimension, on the
 * lines given by (o1, o2) ranges.  lines_full = 1 covers every padded line (the
 * reference fills whole planes); 0 covers interior lines only. */

__global__ void __launch_bounds__(256) k_fill_line(FillArgs a)
{
    const DevState* st = a.st;
    if (st->err) return;
    const DevGrid& g = a.g;
    const int N = a.dim == 0 ? g.Nx : a.dim == 1 ? g.Ny : g.Nz;
    const int H = a.dim == 0 ? g.Hx : a.dim == 1 ? g.Hy : g.Hz;
    const int topo = a.dim == 0 ? g.tx : a.dim == 1 ? g.ty : g.tz;
    const long long stride = a.dim == 0 ? 1 : a.dim == 1 ? g.Px : (long long)g.Px * g.Py;
    /* the two other dimensions */
    const int P1 = a.dim == 0 ? g.Py : g.Px, P2 = a.dim == 2 ? g.Py : g.Pz;
    const int N1 = a.dim == 0 ? g.Ny : g.Nx, N2 = a.dim == 2 ? g.Ny : g.Nz;
    const int H1 = a.dim == 0 ? g.Hy : g.Hx, H2 = a.dim == 2 ? g.Hy : g.Hz;
    const long long n1 = a.lines_full ? P1 : N1, n2 = a.lines_full ? P2 : N2;
    const long long nlines = n1 * n2 * a.nf;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nlines; t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t / (n1 * n2));
        const long long l = t % (n1 * n2);
        const int o1 = (int)(l % n1) + 1 + (a.lines_full ? 0 : H1);
        const int o2 = (int)(l / n1) + 1 + (a.lines_full ? 0 : H2);
        long long base;    /* index of position 1 along the line */
        if (a.dim == 0) base = gidx(g, 1, o1, o2);
        else if (a.dim == 1) base = gidx(g, o1, 1, o2);
        else base = gidx(g, o1, o2, 1);
        float* c = a.f[q] ? a.f[q] : st->tr[st->tcur ^ 1][q];
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


__global__ void __launch_bounds__(256) k_final(FinalArgs a)
{
    const DevState* st = a.st;
    if (st->err) return;
    const DevGrid& g = a.g;
    const long long nint = (long long)g.Nx * g.Ny * g.Nz;
    const float* const* C0 = st->tr[st->tcur];
    float* const* OUT = st->tr[st->tcur ^ 1];
    const long long sz = (long long)g.Px * g.Py, sy = g.Px;
    const float dt = a.dt;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nint; t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.Nx) + 1 + g.Hx;
        const int j = (int)((t / g.Nx) % g.Ny) + 1 + g.Hy;
        const int k = (int)(t / ((long long)g.Nx * g.Ny)) + 1 + g.Hz;
        const long long c = gidx(g, i, j, k);
        const float vol = g.dx * g.dy * dzF_at(g, k);
        /* z sweep faces k and k+1 */
        const float Az = g.dx * g.dy;
        const float w0 = a.w[c], w1 = a.w[c + sz];
        const float Cf0 = fabsf(w0 * dt / dzC_at(g, k)), Cf1 = fabsf(w1 * dt / dzC_at(g, k + 1));
        const float T0 = Az * w0, T1 = Az * w1;
        const float a0 = (2.0f - Cf0) * (1.0f - Cf0) / 6.0f, b0 = (1.0f - Cf0 * Cf0) / 6.0f;
        const float a1 = (2.0f - Cf1) * (1.0f - Cf1) / 6.0f, b1 = (1.0f - Cf1 * Cf1) / 6.0f;
        const bool last_bounded = (g.tz == PI_BOUNDED) && (k == g.Nz + g.Hz);
        const float dT = lmask(g, c, c + sz, T0 - T1);
        const float coef = dt / vol;
        /* tracer_temp = max(0, c) for the forcing (tracer_forcings.jl:3-5) */
        float tt[PI_NTRACERS];
        for (int q = 0; q < PI_NTRACERS; ++q) tt[q] = jl_maxf(0.0f, C0[q][c]);
        const float Ax = g.dy * dzF_at(g, k), Ay = g.dx * dzF_at(g, k);
        for (int q = 0; q < PI_NTRACERS; ++q) {
            const float* src = a.prev[q];
            const float* co = C0[q];
            const float F0 = flux_z(g, src, c, sz, Cf0, a0, b0, T0);
            const float F1 = last_bounded ? 0.0f : flux_z(g, src, c + sz, sz, Cf1, a1, b1, T1);
            const float dG = lmask(g, c, c + sz, F0 - F1);
            const float temp3 = src[c] - coef * (dG - co[c] * dT);
            /* Gcs = tracer_temp - tracer (calc_tracer_tendency!) */
            float G = temp3 - co[c];
            /* tracer_diffusion! (Advection/tracer_diffusion.jl:1-30) on the original c */
            {
                const float fxa = a.kh * Ax * (lmask(g, c - 1, c, co[c] - co[c - 1]) / g.dx);
                const float fxb = a.kh * Ax * (lmask(g, c, c + 1, co[c + 1] - co[c]) / g.dx);
                const float fya = a.kh * Ay * (lmask(g, c - sy, c, co[c] - co[c - sy]) / g.dy);
                const float fyb = a.kh * Ay * (lmask(g, c, c + sy, co[c + sy] - co[c]) / g.dy);
                const float fza = a.kv * Az * (lmask(g, c - sz, c, co[c - sz] - co[c]) / dzF_at(g, k));
                const float fzb = a.kv * Az * (lmask(g, c, c + sz, co[c] - co[c + sz]) / dzF_at(g, k + 1));
                const float ddx = lmask(g, c, c + 1, fxb - fxa);
                const float ddy = lmask(g, c, c + sy, fyb - fya);
                const float ddz = lmask(g, c, c + sz, fza - fzb);
                const float lap = 1.0f / vol * (ddx + ddy + ddz);
                G = G + lap * dt;
            }
            /* tracer_sinking! (Advection/tracer_sinking.jl:1-43) */
            if (q == PI_PFe_inorg || q == PI_Dust || q == PI_PFe_bio) {
                const float ws = q == PI_PFe_bio ? a.wsink_org : a.wsink_inorg;
                const float fa = (k - 1 <= g.Hz) ? 0.0f : ws * co[c - sz];
                const float fb = (k >= g.Nz + g.Hz) ? 0.0f : ws * co[c];
                G = G + (fa - fb) / dzC_at(g, k - g.Hz) * dt;
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
                    const float ex = jl_maxf((DFe - a.lig), 0.0f);
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
                    const float ex = jl_maxf((DFe - a.lig), 0.0f);
                    G = G - tt[PI_PFe_inorg] * a.kdiss * dt +
                        (a.lmin + a.ldust * tt[PI_Dust]) * DFe * a.DFeFrac * dt +
                        ex * DFe * a.DFeFrac * a.lFe * dt;
                    break;
                }
                default: break;
            }
            /* apply_tendency_kernel! (tracer_update.jl:6-20) */
            const float pk = a.plk[q] ? a.plk[q][c] : 0.0f;
            OUT[q][c] = co[c] + (G + pk / vol);
        }
    }
}

/* ------------------------------------------------------------- launchers */
void launch_sweep(hipStream_t s, int blocks, const SweepArgs& a) { hipLaunchKernelGGL(k_sweep, dim3(blocks), dim3(256), 0, s, a); }
void launch_fill(hipStream_t s, int blocks, const FillArgs& a) { hipLaunchKernelGGL(k_fill_line, dim3(blocks), dim3(256), 0, s, a); }
void launch_final(hipStream_t s, int blocks, const FinalArgs& a) { hipLaunchKernelGGL(k_final, dim3(blocks), dim3(256), 0, s, a); }
