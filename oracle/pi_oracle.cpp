/*
 * pi_oracle.cpp - TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * A CPU restatement, kernel by kernel and in the same floating-point operation
 * order, of PlanktonIndividuals.jl v0.7.5's CPU() time step for QuotaMode,
 * CarbonMode, MacroMolecularMode and IronEnergyMode, with abiotic particles
 * (without their interaction with phytoplankton), as run by `julia -t 1` (KernelAbstractions' CPU
 * backend with one thread visits every ndrange index in ascending order, so
 * every `@atomic` accumulation happens in slot order, kernel after kernel).
 *
 * Every function cites the reference file:line it restates.  Julia semantics
 * that matter for bits: Bool*x (jl_mask), NaN-propagating min/max, isless,
 * unsafe_trunc, Float32 * Int promotions, broadcast unaliasing of halo copies.
 * Transcendentals and the random stream come from include/pi_math.h (the
 * stream replaces the reference's entropy-seeded MersenneTwister).
 *
 * Deliberate, documented deviations (none changes a Float32 value of an
 * active individual or of the interior tracer fields):
 *  - ΔzC(...) in particle_diffusion.jl:4-6 reads dzC[zi+2Hz] which is one
 *    element past the array when a particle sits exactly at z == Nz; the
 *    restatement clamps that read to the last element.
 *  - Individual rate fields (PS, VDOC, ...) are kept internally but only the
 *    state fields listed in pi_hip.h are downloadable.
 */
#include "pi_oracle.h"
#include "../include/pi_math.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static thread_local std::string g_err;
static pi_status fail(pi_status s, const std::string& msg) { g_err = msg; return s; }

/* Threads of the element-wise loops (orc_set_threads; 1 by default).  Only
 * loops whose iterations write disjoint elements run in parallel; every
 * accumulation (the per-cell @atomic sums in slot order, the FP32 cumsum and
 * dot of the division step) stays serial, so the results are bit-identical for
 * any thread count. */
static int g_threads = 1;
#define OMP_FOR _Pragma("omp parallel for schedule(static) num_threads(g_threads) if(g_threads > 1)")

template <typename FT> struct N_;
template <> struct N_<float> {
    static float max(float a, float b) { return jl_maxf(a, b); }
    static float min(float a, float b) { return jl_minf(a, b); }
    static int isless(float a, float b) { return jl_islessf(a, b); }
    static float mask(int b, float x) { return jl_maskf(b, x); }
    static float mod1(float x) { return jl_mod1f(x); }
    static float exp(float x) { return pi_expf(x); }
    static float pow2(float x) { return pi_pow2f(x); }
    static float pow4(float x) { return pi_pow4f(x); }
    static float sqrt(float x) { return std::sqrt(x); }
    static float rem(float x, float y) { return std::fmod(x, y); }
    static int isequal(float a, float b) { return jl_isequalf(a, b); }
    static float pow(float x, float y) { return pi_powf(x, y); }
};
template <> struct N_<double> {
    static double max(double a, double b) { return jl_maxd(a, b); }
    static double min(double a, double b) { return jl_mind(a, b); }
    static int isless(double a, double b) { return jl_islessd(a, b); }
    static double mask(int b, double x) { return jl_maskd(b, x); }
    static double mod1(double x) { return jl_mod1d(x); }
    static double exp(double x) { return pi_exp_f64(x); }
    static double pow2(double x) { return pi_pow2d(x); }
    static double pow4(double x) { return pi_pow4d(x); }
    static double sqrt(double x) { return std::sqrt(x); }
    static double rem(double x, double y) { return std::fmod(x, y); }
    static int isequal(double a, double b) { return jl_isequald(a, b); }
    static double pow(double x, double y) {
        if (x != x || x < 0.0) return NAN;
        if (x == 0.0) return 0.0;
        return pi_exp_f64(y * pi_log_f64(x));
    }
};

/* Float32-rounded literals (Julia `0.5f0` etc.), widened in FP64 mode. */
#define L(v) ((FT)(float)(v))

struct orc_model { virtual ~orc_model() {} int precision = 32; };

template <typename FT>
struct Oracle : orc_model {
    using M = N_<FT>;
    typedef std::vector<FT> V;

    /* ---- grid (Grids/rectilinear_grid.jl:1-25) ---- */
    int Nx, Ny, Nz, Hx, Hy, Hz, Px, Py, Pz;
    int64_t ncell;
    int tx, ty, tz;
    FT dx, dy;
    V dzC, dzF;                     /* 1-based: dzC[k-1] */
    std::vector<uint8_t> lm;        /* padded landmask */
    int nsp;
    int64_t cap;
    std::vector<std::array<FT, PI_NPARAMS>> P;
    std::vector<int> dvid_type;
    FT B[PI_NBGC];
    uint64_t seed;
    int64_t iteration = 0;
    int64_t ind_steps = 0;          /* Σ over steps of the individuals stepped (alive at the gather) */
    FT t = 0;
    std::vector<int> diag_tr;
    /* flux boundary conditions per tracer and face (set_bc!, boundary_conditions.jl:12-29):
     * getbc's kind, the Number, the face array(s) (column-major, face size x nt) */
    int bck[PI_NTRACERS][PI_NBC] = {};
    FT bcv[PI_NTRACERS][PI_NBC] = {};
    V bca[PI_NTRACERS][PI_NBC];
    int64_t bcnt[PI_NTRACERS][PI_NBC] = {};
    bool failed = false;
    int mode = PI_MODE_QUOTA;       /* PlanktonModel(...; mode) */

    struct Plank {
        V x, y, z, Sz, Bm, Cq, Nq, Pq, Chl, gen, age;    /* CarbonMode: Bd in Cq */
        V PS, VDOC, VNH4, VNO3, VPO4, rho, resp, BS, exu, graz, mort, dvid;
        V TD, RP;                                     /* CarbonMode thermal damage and repair */
        /* MacroMolecularMode: RNA, and the synthesis rates; PRO, DNA, CH, NST,
         * PST are the columns Bm, Sz, Cq, Nq, Pq */
        V RNA, S_PRO, S_DNA, S_RNA;
        /* IronEnergyMode: CH, qNH4, qP are the columns Cq, Nq, Pq; the other
         * pools and rates (allocated for IronEnergyMode models only) */
        V qNO3, qFe, qFePS, qFeNR, qFeNF, tdark, exEn, ptc;
        V VFe, RS, ERS, CF, ECF, NR, ENR, NF, ENF, ST2PS, PS2ST, ST2NR, NR2ST, ST2NF, NF2ST;
        std::vector<int64_t> xi, yi, zi, idx;
        std::vector<uint8_t> ac;
    };
    std::vector<Plank> sp;

    V tr[PI_NTRACERS], Gcs[PI_NTRACERS], ttemp[PI_NTRACERS], plk[PI_NTRACERS];
    V vel0[3], velh[3], vel1[3];
    V PARF, temp, flux_sink, par, par0, Chl, pop;
    V diagPAR;
    std::vector<V> diagTr;
    std::vector<std::array<V, 5>> diagSp;   /* num, graz, mort, dvid, ptc */
    /* abiotic particles (Individuals/Abiotic): per species the x, y, z, ac and
     * xi, yi, zi columns (in a Plank), parameters, diags.abiotics.saN.num and
     * the particle flux boundary conditions */
    int nsa = 0;
    /* palat.intac / palat.release pairs (species, abiotic species), max_candidates
     * and timestepper.intac (max_candidates x capacity, models.jl:100-104) */
    std::vector<std::array<int, 2>> pint, prel;
    int mc = 25;
    std::vector<int64_t> intac;
    std::vector<std::array<FT, PI_NAPARAMS>> AP;
    std::vector<Plank> ab;
    std::vector<V> abNum;
    int abck[PI_MAXSA][PI_NBC] = {};
    FT abcv[PI_MAXSA][PI_NBC] = {};
    V abca[PI_MAXSA][PI_NBC];
    int64_t abcnt[PI_MAXSA][PI_NBC] = {};
    /* timestepper scratch (timestepper.jl:40-55) */
    V rx, ry, rz, vx, vy, vz, u1, v1, w1, u2, v2, w2;
    V tNH4, tNO3, tPO4, tDOC, tDFe, tpar, tT, tpop, tdpar, idc;
    std::vector<int64_t> idc_int;

    inline int64_t I(int i, int j, int k) const {
        return (int64_t)(i - 1) + (int64_t)Px * ((int64_t)(j - 1) + (int64_t)Py * (int64_t)(k - 1));
    }
    inline bool LM(int i, int j, int k) const { return lm[I(i, j, k)] != 0; }

    /* ---- metrics (Grids/areas_volumes_spacings.jl:9-22) ---- */
    inline FT dzC_at(int64_t k) const {   /* g.dzC[k], 1-based; clamp: see header */
        if (k < 1) k = 1;
        if (k > Pz) k = Pz;
        return dzC[k - 1];
    }
    inline FT dzF_at(int k) const { return dzF[k - 1]; }
    /* LatLonGrid (Grids/lat_lon_grid.jl:1-37): the spacings, areas and volumes
     * are arrays with halos, (Px, Py) or (Px, Py, Pz); RectilinearGrid: Δx, Δy
     * and dzC, dzF (Grids/areas_volumes_spacings.jl:9-38 for both) */
    bool gm = false;
    V mdxC, mdyC, mdxF, mdyF, mAz, mAx, mAy, mVol;
    inline int64_t I2(int i, int j) const { return (int64_t)(i - 1) + (int64_t)Px * (j - 1); }
    inline FT DxC(int i, int j) const { return gm ? mdxC[I2(i, j)] : dx; }
    inline FT DyC(int i, int j) const { return gm ? mdyC[I2(i, j)] : dy; }
    inline FT DxF(int i, int j) const { return gm ? mdxF[I2(i, j)] : dx; }
    inline FT DyF(int i, int j) const { return gm ? mdyF[I2(i, j)] : dy; }
    inline FT Ax(int i, int j, int k) const { return gm ? mAx[I(i, j, k)] : dy * dzF_at(k); }
    inline FT Ay(int i, int j, int k) const { return gm ? mAy[I(i, j, k)] : dx * dzF_at(k); }
    inline FT Az(int i, int j) const { return gm ? mAz[I2(i, j)] : dx * dy; }
    inline FT vol(int i, int j, int k) const { return gm ? mVol[I(i, j, k)] : dx * dy * dzF_at(k); }
    /* a padded index kept inside the array (the random walk's stale index + H) */
    static int clampi(int64_t i, int P) { return (int)(i < 1 ? 1 : i > P ? P : i); }

    /* =============================================================== setup */
    pi_status init(const pi_model_desc* d) {
        const pi_grid_desc& g = d->grid;
        Nx = g.Nx; Ny = g.Ny; Nz = g.Nz; Hx = g.Hx; Hy = g.Hy; Hz = g.Hz;
        tx = g.topo_x; ty = g.topo_y; tz = g.topo_z;
        if (Nx < 1 || Ny < 1 || Nz < 1 || Hx < 2 || Hy < 2 || Hz < 2)
            return fail(PI_ERR_ARG, "grid: N >= 1 and halo >= 2 required");
        if (tz != PI_BOUNDED) return fail(PI_ERR_ARG, "grid: z topology must be Bounded (tracer_sinking.jl)");
        Px = Nx + 2 * Hx; Py = Ny + 2 * Hy; Pz = Nz + 2 * Hz;
        ncell = (int64_t)Px * Py * Pz;
        dx = (FT)g.dx; dy = (FT)g.dy;
        if (d->metrics) {
            const pi_metrics& mt = *d->metrics;
            if (!mt.dxC || !mt.dyC || !mt.dxF || !mt.dyF || !mt.Az || !mt.Ax || !mt.Ay || !mt.Vol)
                return fail(PI_ERR_ARG, "grid metrics: every array is required");
            gm = true;
            const int64_t n2 = (int64_t)Px * Py;
            auto cp = [&](V& dst, const float* src, int64_t n) { dst.resize(n); for (int64_t a = 0; a < n; ++a) dst[a] = (FT)src[a]; };
            cp(mdxC, mt.dxC, n2); cp(mdyC, mt.dyC, n2); cp(mdxF, mt.dxF, n2); cp(mdyF, mt.dyF, n2); cp(mAz, mt.Az, n2);
            cp(mAx, mt.Ax, ncell); cp(mAy, mt.Ay, ncell); cp(mVol, mt.Vol, ncell);
        }
        dzC.resize(Pz); dzF.resize(Pz);
        for (int k = 0; k < Pz; ++k) { dzC[k] = (FT)g.dzC[k]; dzF[k] = (FT)g.dzF[k]; }
        lm.assign(ncell, 0);
        if (g.landmask) {
            for (int64_t c = 0; c < ncell; ++c) lm[c] = g.landmask[c] ? 1 : 0;
        } else {
            /* landmask_validation with landmask = ones (Grids/utils.jl:92-129) */
            std::vector<float> tmp(ncell, 0.0f);
            for (int k = Hz + 1; k <= Hz + Nz; ++k)
                for (int j = Hy + 1; j <= Hy + Ny; ++j)
                    for (int i = Hx + 1; i <= Hx + Nx; ++i) tmp[I(i, j, k)] = 1.0f;
            orc_fill_halo_lm(tmp);
            for (int64_t c = 0; c < ncell; ++c) lm[c] = tmp[c] != 0.0f;
        }
        nsp = d->n_species; cap = d->capacity;
        if (nsp < 1 || cap < 1) return fail(PI_ERR_ARG, "n_species and capacity must be >= 1");
        mode = d->mode;
        if (mode < PI_MODE_QUOTA || mode > PI_MODE_IRONENERGY) return fail(PI_ERR_UNSUPPORTED, "physiology mode");
        P.resize(nsp); dvid_type.resize(nsp);
        for (int s = 0; s < nsp; ++s) {
            for (int q = 0; q < PI_NPARAMS; ++q) P[s][q] = (FT)d->params[s * PI_NPARAMS + q];
            const int dt_idx = mode == PI_MODE_CARBON ? (int)PI_C_dvid_type
                             : mode == PI_MODE_IRONENERGY ? (int)PI_I_dvid_type : (int)PI_P_dvid_type;
            dvid_type[s] = mode == PI_MODE_MACROMOLECULAR ? 1 : (int)d->params[s * PI_NPARAMS + dt_idx];
            if (dvid_type[s] < 1 || dvid_type[s] > 3)
                return fail(PI_ERR_ARG, "Wrong cell division type, must be in 1 to 3");
        }
        for (int q = 0; q < PI_NBGC; ++q) B[q] = (FT)d->bgc[q];
        seed = d->rng_seed;
        for (int q = 0; q < d->n_diag_tracers; ++q) diag_tr.push_back(d->diag_tracers[q]);

        sp.resize(nsp);
        for (auto& p : sp) {
            for (V* f : {&p.x, &p.y, &p.z, &p.Sz, &p.Bm, &p.Cq, &p.Nq, &p.Pq, &p.Chl, &p.gen, &p.age,
                         &p.PS, &p.VDOC, &p.VNH4, &p.VNO3, &p.VPO4, &p.rho, &p.resp, &p.BS, &p.exu,
                         &p.graz, &p.mort, &p.dvid, &p.TD, &p.RP, &p.RNA, &p.S_PRO, &p.S_DNA, &p.S_RNA})
                f->assign(cap, (FT)0);
            p.xi.assign(cap, 0); p.yi.assign(cap, 0); p.zi.assign(cap, 0); p.idx.assign(cap, 0);
            p.ac.assign(cap, 0);
            if (mode == PI_MODE_IRONENERGY)
                for (V* f : {&p.qNO3, &p.qFe, &p.qFePS, &p.qFeNR, &p.qFeNF, &p.tdark, &p.exEn, &p.ptc, &p.VFe, &p.RS,
                             &p.ERS, &p.CF, &p.ECF, &p.NR, &p.ENR, &p.NF, &p.ENF, &p.ST2PS, &p.PS2ST, &p.ST2NR,
                             &p.NR2ST, &p.ST2NF, &p.NF2ST})
                    f->assign(cap, (FT)0);
        }
        for (int q = 0; q < PI_NTRACERS; ++q) {
            tr[q].assign(ncell, 0); Gcs[q].assign(ncell, 0); ttemp[q].assign(ncell, 0); plk[q].assign(ncell, 0);
        }
        for (int q = 0; q < 3; ++q) { vel0[q].assign(ncell, 0); velh[q].assign(ncell, 0); vel1[q].assign(ncell, 0); }
        PARF.assign((int64_t)Nx * Ny, 0);
        for (V* f : {&temp, &flux_sink, &par, &par0, &Chl, &pop, &diagPAR}) f->assign(ncell, 0);
        diagTr.assign(diag_tr.size(), V(ncell, 0));
        diagSp.resize(nsp);
        for (auto& a : diagSp) for (auto& v : a) v.assign(ncell, 0);
        nsa = d->n_abiotic;
        if (nsa < 0 || nsa > PI_MAXSA) return fail(PI_ERR_ARG, "n_abiotic out of range");
        if (d->n_interactions && mode != PI_MODE_IRONENERGY)
            return fail(PI_ERR_ARG, "phytoplankton-particle interaction needs IronEnergyMode (max_ptc, ptc)");
        for (int q = 0; q < d->n_interactions; ++q) pint.push_back({d->interactions[2 * q], d->interactions[2 * q + 1]});
        for (int q = 0; q < d->n_releases; ++q) prel.push_back({d->releases[2 * q], d->releases[2 * q + 1]});
        mc = d->max_candidates > 0 ? d->max_candidates : 25;
        if (!pint.empty()) intac.assign((size_t)mc * cap, 0);
        AP.resize(nsa); ab.resize(nsa); abNum.assign(nsa, V(ncell, 0));
        for (int a = 0; a < nsa; ++a) {
            for (int q = 0; q < PI_NAPARAMS; ++q) AP[a][q] = (FT)d->abiotic_params[a * PI_NAPARAMS + q];
            Plank& p = ab[a];
            for (V* f : {&p.x, &p.y, &p.z}) f->assign(cap, (FT)0);
            p.xi.assign(cap, 0); p.yi.assign(cap, 0); p.zi.assign(cap, 0);
            p.ac.assign(cap, 0);
        }
        for (V* f : {&rx, &ry, &rz, &vx, &vy, &vz, &u1, &v1, &w1, &u2, &v2, &w2, &tNH4, &tNO3, &tPO4,
                     &tDOC, &tDFe, &tpar, &tT, &tpop, &tdpar, &idc})
            f->assign(cap, 0);
        idc_int.assign(cap, 0);
        return PI_OK;
    }

    void orc_fill_halo_lm(std::vector<float>& c) {
        /* Grids/utils.jl:113-127: same copies as fill_halo_tracer for Bool arrays */
        fill_halo_generic(c.data(), tx, ty);
    }

    /* ================================================== halo regions ====
     * Biogeochemistry/HaloRegions/halo_regions.jl:1-35.  Each `@views @.`
     * statement copies its source first (broadcast unaliasing), hence the
     * snapshot buffers. */
    template <typename T>
    void planes_copy(T* c, int dim, int d0, int s0, int n) {   /* dst[d0..d0+n) <- src[s0..s0+n) */
        std::vector<T> buf;
        int64_t np = plane_size(dim);
        buf.resize((size_t)(np * n));
        for (int a = 0; a < n; ++a) plane_get(c, dim, s0 + a, buf.data() + a * np);
        for (int a = 0; a < n; ++a) plane_set(c, dim, d0 + a, buf.data() + a * np);
    }
    template <typename T>
    void planes_bcast(T* c, int dim, int d0, int n, int s) {   /* dst[d0..d0+n) <- src plane s */
        std::vector<T> buf((size_t)plane_size(dim));
        plane_get(c, dim, s, buf.data());
        for (int a = 0; a < n; ++a) plane_set(c, dim, d0 + a, buf.data());
    }
    template <typename T>
    void planes_zero(T* c, int dim, int d0, int n) {
        std::vector<T> buf((size_t)plane_size(dim), (T)0);
        for (int a = 0; a < n; ++a) plane_set(c, dim, d0 + a, buf.data());
    }
    int64_t plane_size(int dim) const {
        return dim == 0 ? (int64_t)Py * Pz : dim == 1 ? (int64_t)Px * Pz : (int64_t)Px * Py;
    }
    template <typename T>
    void plane_get(const T* c, int dim, int p, T* out) const {
        int64_t n = 0;
        if (dim == 0) { for (int k = 1; k <= Pz; ++k) for (int j = 1; j <= Py; ++j) out[n++] = c[I(p, j, k)]; }
        else if (dim == 1) { for (int k = 1; k <= Pz; ++k) for (int i = 1; i <= Px; ++i) out[n++] = c[I(i, p, k)]; }
        else { for (int j = 1; j <= Py; ++j) for (int i = 1; i <= Px; ++i) out[n++] = c[I(i, j, p)]; }
    }
    template <typename T>
    void plane_set(T* c, int dim, int p, const T* in) const {
        int64_t n = 0;
        if (dim == 0) { for (int k = 1; k <= Pz; ++k) for (int j = 1; j <= Py; ++j) c[I(p, j, k)] = in[n++]; }
        else if (dim == 1) { for (int k = 1; k <= Pz; ++k) for (int i = 1; i <= Px; ++i) c[I(i, p, k)] = in[n++]; }
        else { for (int j = 1; j <= Py; ++j) for (int i = 1; i <= Px; ++i) c[I(i, j, p)] = in[n++]; }
    }
    int Nd(int dim) const { return dim == 0 ? Nx : dim == 1 ? Ny : Nz; }
    int Hd(int dim) const { return dim == 0 ? Hx : dim == 1 ? Hy : Hz; }
    int Td(int dim) const { return dim == 0 ? tx : dim == 1 ? ty : tz; }
    /* fill_halo_west!/south!/top! (halo_regions.jl:2-4,10-12) */
    template <typename T> void fill_lo(T* c, int dim) {
        int N = Nd(dim), H = Hd(dim);
        if (Td(dim) == PI_PERIODIC) planes_copy(c, dim, 1, N + 1, H);
        else planes_bcast(c, dim, 1, H, H + 1);
    }
    /* fill_halo_east!/north!/bottom! (halo_regions.jl:6-8,14-16) */
    template <typename T> void fill_hi(T* c, int dim) {
        int N = Nd(dim), H = Hd(dim);
        if (Td(dim) == PI_PERIODIC) planes_copy(c, dim, N + H + 1, 1 + H, H);
        else planes_bcast(c, dim, N + H + 1, H, N + H);
    }
    /* fill_halo_*_vel! (halo_regions.jl:18-20,26-28) */
    template <typename T> void fill_hi_vel(T* c, int dim) {
        int N = Nd(dim), H = Hd(dim);
        if (Td(dim) == PI_PERIODIC) fill_hi(c, dim);
        else planes_bcast(c, dim, N + H + 2, H - 1, N + H + 1);
    }
    /* fill_halo_*_Gc! (halo_regions.jl:22-24,30-32) */
    template <typename T> void fill_hi_Gc(T* c, int dim) {
        int N = Nd(dim), H = Hd(dim);
        if (Td(dim) == PI_PERIODIC) fill_hi(c, dim);
        else planes_zero(c, dim, N + H + 1, H);
    }
    /* fill_halo_top_flux_sink! / fill_halo_bottom_flux_sink! (halo_regions.jl:34-35):
     * zero planes 1:H and N+H:N+2H (the last interior level included) */
    template <typename T> void fill_top_flux_sink(T* c) { planes_zero(c, 2, 1, Hz); }
    template <typename T> void fill_bottom_flux_sink(T* c) { planes_zero(c, 2, Nz + Hz, Hz + 1); }
    template <typename T> void fill_halo_generic(T* c, int tx_, int ty_) {
        (void)tx_; (void)ty_;
        fill_lo(c, 0); fill_hi(c, 0); fill_lo(c, 1); fill_hi(c, 1); fill_lo(c, 2); fill_hi(c, 2);
    }
    /* fill_halo_tracer! (halo_regions.jl:37-47) */
    void fill_halo_tracer(V* f) {
        for (int q = 0; q < PI_NTRACERS; ++q) fill_halo_generic(f[q].data(), tx, ty);
    }
    /* fill_halo_Gcs! (halo_regions.jl:49-56) */
    void fill_halo_Gcs(V* f) {
        for (int q = 0; q < PI_NTRACERS; ++q) { fill_hi_Gc(f[q].data(), 0); fill_hi_Gc(f[q].data(), 1); fill_hi_Gc(f[q].data(), 2); }
    }
    /* fill_halo_u!/v!/w! (halo_regions.jl:64-99) */
    void fill_halo_vel(V* v) {
        FT* u = v[0].data(); FT* vv = v[1].data(); FT* w = v[2].data();
        fill_hi_vel(u, 0); fill_lo(u, 0); fill_lo(u, 1); fill_hi(u, 1); fill_lo(u, 2); fill_hi(u, 2);
        fill_hi_vel(vv, 1); fill_lo(vv, 0); fill_hi(vv, 0); fill_lo(vv, 1); fill_lo(vv, 2); fill_hi(vv, 2);
        fill_hi_vel(w, 2); fill_lo(w, 0); fill_hi(w, 0); fill_lo(w, 1); fill_hi(w, 1); fill_lo(w, 2);
    }

    /* ================================================== inputs ==========
     * vel_copy! / copy_interior! (Simulation/utils.jl:5-34) */
    void set_velocity(const float* u, const float* v, const float* w) {
        int nxu = Nx + (tx == PI_BOUNDED), nyv = Ny + (ty == PI_BOUNDED), nzw = Nz + (tz == PI_BOUNDED);
        for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < nxu; ++i)
            vel1[0][I(Hx + 1 + i, Hy + 1 + j, Hz + 1 + k)] = (FT)u[i + (int64_t)nxu * (j + (int64_t)Ny * k)];
        for (int k = 0; k < Nz; ++k) for (int j = 0; j < nyv; ++j) for (int i = 0; i < Nx; ++i)
            vel1[1][I(Hx + 1 + i, Hy + 1 + j, Hz + 1 + k)] = (FT)v[i + (int64_t)Nx * (j + (int64_t)nyv * k)];
        for (int k = 0; k < nzw; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
            vel1[2][I(Hx + 1 + i, Hy + 1 + j, Hz + 1 + k)] = (FT)w[i + (int64_t)Nx * (j + (int64_t)Ny * k)];
        fill_halo_vel(vel1);
    }
    void set_temperature(const float* tt) {
        for (int k = 0; k < Nz; ++k) for (int j = 0; j < Ny; ++j) for (int i = 0; i < Nx; ++i)
            temp[I(Hx + 1 + i, Hy + 1 + j, Hz + 1 + k)] = (FT)tt[i + (int64_t)Nx * (j + (int64_t)Ny * k)];
    }

    /* ================================================== particles =======
     * Individuals/Advection/interpolation.jl:1-83 */
    FT itpl(const FT* f, FT x, FT y, FT z, bool ac, int comp) const {
        /* comp 0: u (x on faces), 1: v (y on faces), 2: w (z on faces) */
        FT xs = comp == 0 ? x + L(1) : x + L(0.5f);
        FT ys = comp == 1 ? y + L(1) : y + L(0.5f);
        FT zs = comp == 2 ? z + L(1) : z + L(0.5f);
        FT xf = M::mask(ac, xs), yf = M::mask(ac, ys), zf = M::mask(ac, zs);
        FT xd = M::mod1(xf), yd = M::mod1(yf), zd = M::mod1(zf);
        int64_t xi = (int64_t)xf + Hx, yi = (int64_t)yf + Hy, zi = (int64_t)zf + Hz;
        const FT one = L(1);
        /* ΔxC / ΔyC / ΔzC of the corner (normalized_trilinear_interpolation_u/v/w) */
        auto D = [&](int64_t i, int64_t j, int64_t k) -> FT {
            return comp == 0 ? DxC((int)i, (int)j) : comp == 1 ? DyC((int)i, (int)j) : dzC_at(k);
        };
        auto F = [&](int64_t i, int64_t j, int64_t k) -> FT { return f[I((int)i, (int)j, (int)k)]; };
        FT s = (one - xd) * (one - yd) * (one - zd) * F(xi, yi, zi) / D(xi, yi, zi);
        s = s + (one - xd) * (one - yd) * zd * F(xi, yi, zi + 1) / D(xi, yi, zi + 1);
        s = s + (one - xd) * yd * (one - zd) * F(xi, yi + 1, zi) / D(xi, yi + 1, zi);
        s = s + (one - xd) * yd * zd * F(xi, yi + 1, zi + 1) / D(xi, yi + 1, zi + 1);
        s = s + xd * (one - yd) * (one - zd) * F(xi + 1, yi, zi) / D(xi + 1, yi, zi);
        s = s + xd * (one - yd) * zd * F(xi + 1, yi, zi + 1) / D(xi + 1, yi, zi + 1);
        s = s + xd * yd * (one - zd) * F(xi + 1, yi + 1, zi) / D(xi + 1, yi + 1, zi);
        s = s + xd * yd * zd * F(xi + 1, yi + 1, zi + 1) / D(xi + 1, yi + 1, zi + 1);
        return s;
    }
    /* vel_interpolate_kernel! (particle_advection_kernels.jl:26-37) */
    void vel_interpolate(V& ut, V& vt, V& wt, const V& x, const V& y, const V& z,
                         const std::vector<uint8_t>& ac, V* vel) {
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            ut[i] = M::mask(ac[i], itpl(vel[0].data(), x[i], y[i], z[i], ac[i], 0));
            vt[i] = M::mask(ac[i], itpl(vel[1].data(), x[i], y[i], z[i], ac[i], 1));
            wt[i] = M::mask(ac[i], itpl(vel[2].data(), x[i], y[i], z[i], ac[i], 2));
        }
    }
    /* calc_coord_kernel! (particle_advection_kernels.jl:40-50) */
    void calc_coord(V& ox, V& oy, V& oz, const V& px, const V& py, const V& pz,
                    const V& u, const V& v, const V& w, const std::vector<uint8_t>& ac, FT dt, FT wt) {
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            ox[i] = px[i] + M::mask(ac[i], wt * u[i] * dt);
            oy[i] = py[i] + M::mask(ac[i], wt * v[i] * dt);
            oz[i] = pz[i] - M::mask(ac[i], wt * w[i] * dt);
        }
    }
    /* particle_boundary_condition (particle_advection_kernels.jl:2-11) */
    static FT pbc(FT x, int N, int topo) {
        const FT xl = (FT)0, xr = (FT)N;
        if (topo == PI_PERIODIC) {
            if (x > xr) return xl + (x - xr);
            if (x < xl) return xr - (xl - x);
            return x;
        }
        if (x > xr) return xr - (x - xr) * L(1);
        if (x < xl) return xl + (xl - x) * L(1);
        return x;
    }
    /* particle_boundaries_kernel! (particle_advection_kernels.jl:13-23) */
    void boundaries(V& x, V& y, V& z, const std::vector<uint8_t>& ac) {
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            x[i] = M::mask(ac[i], pbc(x[i], Nx, tx));
            y[i] = M::mask(ac[i], pbc(y[i], Ny, ty));
            z[i] = M::mask(ac[i], pbc(z[i], Nz, tz));
        }
    }
    /* particle_advection! RK4 (Individuals/Advection/particle_advection.jl:2-56) */
    void particle_advection(Plank& p, FT dt) {
        vel_interpolate(u1, v1, w1, p.x, p.y, p.z, p.ac, vel0);
        u2 = u1; v2 = v1; w2 = w1;
        calc_coord(vx, vy, vz, p.x, p.y, p.z, u1, v1, w1, p.ac, dt, L(0.5f));
        boundaries(vx, vy, vz, p.ac);
        vel_interpolate(u1, v1, w1, vx, vy, vz, p.ac, velh);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) { u2[i] = u2[i] + u1[i] * L(2); v2[i] = v2[i] + v1[i] * L(2); w2[i] = w2[i] + w1[i] * L(2); }
        calc_coord(vx, vy, vz, p.x, p.y, p.z, u1, v1, w1, p.ac, dt, L(0.5f));
        boundaries(vx, vy, vz, p.ac);
        vel_interpolate(u1, v1, w1, vx, vy, vz, p.ac, velh);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) { u2[i] = u2[i] + u1[i] * L(2); v2[i] = v2[i] + v1[i] * L(2); w2[i] = w2[i] + w1[i] * L(2); }
        calc_coord(vx, vy, vz, p.x, p.y, p.z, u1, v1, w1, p.ac, dt, L(1));
        boundaries(vx, vy, vz, p.ac);
        vel_interpolate(u1, v1, w1, vx, vy, vz, p.ac, vel1);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) { u2[i] = u2[i] + u1[i]; v2[i] = v2[i] + v1[i]; w2[i] = w2[i] + w1[i]; }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) { u2[i] = u2[i] / L(6); v2[i] = v2[i] / L(6); w2[i] = w2[i] / L(6); }
        calc_coord(p.x, p.y, p.z, p.x, p.y, p.z, u2, v2, w2, p.ac, dt, L(1));
        boundaries(p.x, p.y, p.z, p.ac);
    }
    /* particle_diffusion! (Individuals/Advection/particle_diffusion.jl:2-33) */
    void particle_diffusion(Plank& p, int s, FT dt) {
        const FT kx = B[PI_B_kappa_hP], ky = B[PI_B_kappa_hP], kz = B[PI_B_kappa_vP];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            float a, b, c;
            pi_uniform3(seed, iteration, s, PI_KIND_DIFFUSION, i, &a, &b, &c);
            rx[i] = (FT)a; ry[i] = (FT)b; rz[i] = (FT)c;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            rx[i] = (rx[i] - L(0.5f)) * L(2);
            ry[i] = (ry[i] - L(0.5f)) * L(2);
            rz[i] = (rz[i] - L(0.5f)) * L(2);
        }
        const FT sx = M::sqrt(kx * dt), sy = M::sqrt(ky * dt), sz = M::sqrt(kz * dt);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_diffusion_kernel! :2-8, stale xi/yi/zi (+H twice) */
            const int ci = clampi(p.xi[i] + Hx, Px), cj = clampi(p.yi[i] + Hy, Py);
            rx[i] = rx[i] * sx / DxC(ci, cj);
            ry[i] = ry[i] * sy / DyC(ci, cj);
            rz[i] = rz[i] * sz / dzC_at(p.zi[i] + Hz);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            p.x[i] = p.x[i] + M::mask(p.ac[i], rx[i]);
            p.y[i] = p.y[i] + M::mask(p.ac[i], ry[i]);
            p.z[i] = p.z[i] + M::mask(p.ac[i], rz[i]);
        }
        boundaries(p.x, p.y, p.z, p.ac);
    }
    /* find_inds_kernel! (Individuals/utils.jl:2-12) */
    void find_inds(Plank& p) {
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            p.xi[i] = (int64_t)M::mask(p.ac[i], p.x[i] + L(1)) + Hx;
            p.yi[i] = (int64_t)M::mask(p.ac[i], p.y[i] + L(1)) + Hy;
            p.zi[i] = (int64_t)M::mask(p.ac[i], p.z[i] + L(1)) + Hz;
        }
    }
    inline int64_t cell_of(const Plank& p, int64_t i) const { return I((int)p.xi[i], (int)p.yi[i], (int)p.zi[i]); }
    /* acc_chl_kernel! / acc_counts_kernel! (Individuals/utils.jl:33-52) */
    void acc_chl(const Plank& p) {
        for (int64_t i = 0; i < cap; ++i) Chl[cell_of(p, i)] += M::mask(p.ac[i], p.Chl[i]);
    }
    void acc_counts(const Plank& p) {
        for (int64_t i = 0; i < cap; ++i) pop[cell_of(p, i)] += (FT)(p.ac[i] ? 1 : 0);
    }
    /* calc_par_kernel! (Individuals/utils.jl:55-68), one call per level (time_step.jl:63-67) */
    void calc_par(int ki) {
        const FT kc = B[PI_B_kc], kw = B[PI_B_kw];
        OMP_FOR
        for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
            int ii = i + Hx, jj = j + Hy, kk = ki + Hz;
            FT atten = (Chl[I(ii, jj, kk)] / vol(ii, jj, kk) * kc + kw) * dzF_at(kk);
            int64_t f = (int64_t)(i - 1) + (int64_t)Nx * (j - 1);
            par[I(ii, jj, kk)] = PARF[f] * (L(1) - M::exp(-atten)) / atten;
            PARF[f] = PARF[f] * M::exp(-atten);
        }
    }
    /* find_NPT_kernel! (Individuals/utils.jl:14-30) */
    void find_NPT(const Plank& p) {
        const FT z0 = L(0);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            int64_t c = cell_of(p, i);
            bool a = p.ac[i];
            tNH4[i] = M::mask(a, M::max(z0, tr[PI_NH4][c]));
            tNO3[i] = M::mask(a, M::max(z0, tr[PI_NO3][c]));
            tPO4[i] = M::mask(a, M::max(z0, tr[PI_PO4][c]));
            tDOC[i] = M::mask(a, M::max(z0, tr[PI_DOC][c]));
            tDFe[i] = M::mask(a, M::max(z0, tr[PI_DFe][c]));
            tpar[i] = M::mask(a, par[c]);
            tdpar[i] = M::mask(a, par[c] - par0[c]);
            tT[i] = M::mask(a, temp[c]);
            tpop[i] = M::mask(a, pop[c]);
        }
    }

    /* ---- shape functions (Individuals/utils.jl:109-121) ---- */
    static FT shape_dec(FT x, FT xmax, FT k, int pw) {
        FT fx = M::max(L(0), M::min(L(1), L(1) - x / xmax));
        FT fp = pw == 2 ? M::pow2(fx) : M::pow4(fx);
        return fp / (k + fp);
    }
    static FT shape_inc(FT x, FT xmax, FT k, int pw) { return L(1) - shape_dec(x, xmax, k, pw); }

    /* ---- QuotaMode growth (Plankton/QuotaMode/growth_kernels.jl:1-46) ---- */
    static FT tempFunc_PS(FT T, const std::array<FT, PI_NPARAMS>& q) {
        FT x = T - q[PI_P_Topt], xmax = q[PI_P_Tmax] - q[PI_P_Topt];
        FT regT = shape_dec(x, xmax, L(4.0e-2f), 4);
        FT k = M::exp(-q[PI_P_Ea] / (L(8.3145f) * (T + L(273.15f)))) * regT;
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_P_Ea] / (L(8.3145f) * (q[PI_P_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }
    static FT tempFunc(FT T, const std::array<FT, PI_NPARAMS>& q) {
        FT k = M::exp(-q[PI_P_Ea] / (L(8.3145f) * (T + L(273.15f))));
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_P_Ea] / (L(8.3145f) * (q[PI_P_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }

    /* ---- losses (Plankton/QuotaMode/consume_loss.jl:17-33, division_death.jl:2-25) ---- */
    void calc_loss(const Plank& p, const V& loss, FT fC, FT fN, FT fP, FT R_NC, FT R_PC) {
        for (int64_t i = 0; i < cap; ++i) {
            int64_t c = cell_of(p, i);
            bool a = p.ac[i];
            plk[PI_DOC][c] += M::mask(a, (p.Bm[i] + p.Cq[i]) * fC) * loss[i];
            plk[PI_DON][c] += M::mask(a, (p.Bm[i] * R_NC + p.Nq[i]) * fN) * loss[i];
            plk[PI_DOP][c] += M::mask(a, (p.Bm[i] * R_PC + p.Pq[i]) * fP) * loss[i];
            plk[PI_POC][c] += M::mask(a, (p.Bm[i] + p.Cq[i]) * (L(1) - fC)) * loss[i];
            plk[PI_PON][c] += M::mask(a, (p.Bm[i] * R_NC + p.Nq[i]) * (L(1) - fN)) * loss[i];
            plk[PI_POP][c] += M::mask(a, (p.Bm[i] * R_PC + p.Pq[i]) * (L(1) - fP)) * loss[i];
        }
    }
    void inactivate(Plank& p, const V& loss) {   /* Individuals/utils.jl:103-105 */
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.ac[i] = p.ac[i] && M::isless(loss[i], L(1));
    }
    /* diags_proc_kernel! (Diagnostics/diagnostics_kernels.jl:4-12) */
    void diags_proc(V& d, const V& proc, const Plank& p) {
        for (int64_t i = 0; i < cap; ++i) d[cell_of(p, i)] += M::mask(p.ac[i], proc[i]);
    }

    /* ---- plankton_update! (Plankton/QuotaMode/plankton_update.jl:1-34) ---- */
    pi_status plankton_update(int s, FT dt) {
        if (mode == PI_MODE_CARBON) return carbon_update(s, dt);
        if (mode == PI_MODE_MACROMOLECULAR) return mm_update(s, dt);
        if (mode == PI_MODE_IRONENERGY) return iron_update(s, dt);
        Plank& p = sp[s];
        const auto& q = P[s];
        diags_proc(diagSp[s][1], p.graz, p);
        calc_loss(p, p.graz, q[PI_P_grazFracC], q[PI_P_grazFracN], q[PI_P_grazFracP], q[PI_P_R_NC], q[PI_P_R_PC]);
        inactivate(p, p.graz);
        diags_proc(diagSp[s][2], p.mort, p);
        calc_loss(p, p.mort, q[PI_P_mortFracC], q[PI_P_mortFracN], q[PI_P_mortFracP], q[PI_P_R_NC], q[PI_P_R_PC]);
        inactivate(p, p.mort);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.dvid[i] = M::mask(p.ac[i], p.dvid[i]);
        diags_proc(diagSp[s][3], p.dvid, p);
        /* dvidnum = dot(dvid, ac); deactive_ind = findall(!ac) */
        FT dvidnum = 0;
        for (int64_t i = 0; i < cap; ++i) dvidnum += M::mask(p.ac[i], p.dvid[i]);
        std::vector<int64_t> free_ind;
        for (int64_t i = 0; i < cap; ++i) if (!p.ac[i]) free_ind.push_back(i + 1);
        if (dvidnum > (FT)free_ind.size()) {
            char buf[160];
            snprintf(buf, sizeof buf, "number of individual exceeds the capacity at timestep %g days",
                     (double)t / 86400.0);
            failed = true;
            return fail(PI_ERR_CAPACITY, buf);
        }
        divide(p, free_ind);
        plankton_growth(s, dt);
        calc_consume(p, dt);
        /* diags_spcs! is a no-op for the default per-species diagnostics */
        for (int64_t i = 0; i < cap; ++i) diagSp[s][0][cell_of(p, i)] += (FT)((p.ac[i] && p.ac[i]) ? 1 : 0);
        return PI_OK;
    }
    /* divide! (QuotaMode/division_death.jl:28-77, Individuals/utils.jl:76-86) */
    void divide(Plank& p, const std::vector<int64_t>& free_ind) {
        FT acc = 0;
        for (int64_t i = 0; i < cap; ++i) { acc = acc + p.dvid[i]; idc[i] = acc; }   /* accumulate!(+) */
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) idc_int[i] = (int64_t)idc[i];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)                                           /* get_tind_kernel! */
            if (p.dvid[i] == L(1)) p.idx[i] = free_ind[(size_t)(idc_int[i] - 1)];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {                                         /* copy_daughter_individuals_kernel! */
            if (p.dvid[i] == L(1) && p.idx[i] != 0) {
                int64_t d = p.idx[i] - 1;
                p.x[d] = p.x[i]; p.y[d] = p.y[i]; p.z[d] = p.z[i];
                p.Sz[d] = p.Sz[i]; p.Bm[d] = p.Bm[i]; p.Cq[d] = p.Cq[i]; p.Nq[d] = p.Nq[i]; p.Pq[d] = p.Pq[i];
                p.Chl[d] = p.Chl[i]; p.gen[d] = p.gen[i]; p.ac[d] = p.ac[i]; p.dvid[d] = p.dvid[i];
            }
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {                                         /* divide_to_half_kernel! */
            FT h = (L(2) - p.dvid[i]) / L(2);
            p.Sz[i] = p.Sz[i] * h; p.Bm[i] = p.Bm[i] * h; p.Cq[i] = p.Cq[i] * h;
            p.Nq[i] = p.Nq[i] * h; p.Pq[i] = p.Pq[i] * h; p.Chl[i] = p.Chl[i] * h;
            p.gen[i] = p.gen[i] + p.dvid[i];
            p.age[i] = p.age[i] * (L(1) - p.dvid[i]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.idx[i] = 0;
    }
    /* plankton_growth! (QuotaMode/plankton_growth.jl:2-24, growth_kernels.jl:48-171) */
    void plankton_growth(int s, FT dt) {
        Plank& p = sp[s];
        const auto& q = P[s];
        const FT e30 = L(1.0e-30f), z0 = L(0), one = L(1);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_inorganic_uptake_kernel! */
            bool a = p.ac[i];
            FT aI = tpar[i] * q[PI_P_alpha] * q[PI_P_Phi];
            FT PCm = q[PI_P_PCmax] * tempFunc_PS(tT[i], q);
            FT PS = PCm * (one - M::exp(-aI / M::max(e30, PCm) * p.Chl[i] / M::max(e30, p.Bm[i]))) * p.Bm[i];
            p.PS[i] = M::mask(a, PS);
            FT Bm = p.Bm[i], Cq = p.Cq[i];
            FT Qn = (p.Nq[i] + Bm * q[PI_P_R_NC]) / M::max(e30, Bm + Cq);
            FT Qp = (p.Pq[i] + Bm * q[PI_P_R_PC]) / M::max(e30, Bm + Cq);
            FT regQN = shape_dec(Qn, q[PI_P_Nqmax], L(1.0e-4f), 4);
            FT regQP = shape_dec(Qp, q[PI_P_Pqmax], L(1.0e-4f), 4);
            FT T = tT[i];
            FT VNH4 = M::mask(a, q[PI_P_VNH4max] * regQN * tNH4[i] / M::max(e30, tNH4[i] + q[PI_P_KsatNH4]) * tempFunc(T, q) * Bm);
            FT VNO3 = M::mask(a, q[PI_P_VNO3max] * regQN * tNO3[i] / M::max(e30, tNO3[i] + q[PI_P_KsatNO3]) * tempFunc(T, q) * Bm);
            FT VPO4 = M::mask(a, q[PI_P_VPO4max] * regQP * tPO4[i] / M::max(e30, tPO4[i] + q[PI_P_KsatPO4]) * tempFunc(T, q) * Bm);
            FT pp = M::max(one, tpop[i]);
            p.VNH4[i] = M::min(VNH4, tNH4[i] / dt / pp);
            p.VNO3[i] = M::min(VNO3, tNO3[i] / dt / pp);
            p.VPO4[i] = M::min(VPO4, tPO4[i] / dt / pp);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_quotas_1_kernel! */
            p.Cq[i] = p.Cq[i] + p.PS[i] * dt;
            p.Pq[i] = p.Pq[i] + p.VPO4[i] * dt;
            p.Nq[i] = p.Nq[i] + (p.VNH4[i] + p.VNO3[i]) * dt;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_organic_uptake_kernel! */
            FT Bm = p.Bm[i], Cq = p.Cq[i];
            FT Qc = Cq / M::max(e30, Bm + Cq);
            FT regQ = shape_dec(Qc, q[PI_P_Cqmax], L(1.0e-4f), 4);
            FT VN = q[PI_P_VDOCmax] * regQ * tDOC[i] / M::max(e30, tDOC[i] + q[PI_P_KsatDOC]) * tempFunc(tT[i], q) * Bm;
            FT VD = M::min(VN, tDOC[i] / dt / M::max(one, tpop[i]));
            p.VDOC[i] = M::mask(p.ac[i], VD);
            p.VDOC[i] = M::mask(M::isless(L(1.0e-2f), p.PS[i] / (p.VDOC[i] + p.PS[i])), p.VDOC[i]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_ρChl_kernel! */
            FT r = p.PS[i] / M::max(e30, p.Bm[i]) * q[PI_P_Chl2N] /
                   M::max(e30, tpar[i] * q[PI_P_alpha] * q[PI_P_Phi] * p.Chl[i] / p.Bm[i]);
            r = M::mask(M::isless(L(1.0e-1f), tpar[i]), r);
            p.rho[i] = M::mask(p.ac[i], r);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_respir_kernel! */
            p.resp[i] = M::mask(p.ac[i], q[PI_P_respir] * p.Bm[i] * tempFunc(tT[i], q));
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_quotas_2_kernel! */
            p.Cq[i] = p.Cq[i] + (p.VDOC[i] - p.resp[i]) * dt;
            p.Bm[i] = p.Bm[i] - M::max(z0, (z0 - p.Cq[i]));
            p.Nq[i] = p.Nq[i] + M::max(z0, (z0 - p.Cq[i])) * q[PI_P_R_NC];
            p.Pq[i] = p.Pq[i] + M::max(z0, (z0 - p.Cq[i])) * q[PI_P_R_PC];
            p.Cq[i] = p.Cq[i] + M::max(z0, (z0 - p.Cq[i]));
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_BS_kernel! */
            FT mn = M::min(M::min(p.Cq[i], p.Nq[i] / q[PI_P_R_NC]), p.Pq[i] / q[PI_P_R_PC]);
            p.BS[i] = M::mask(p.ac[i], mn * q[PI_P_k_mtb]);
            FT mn2 = M::min(M::min(p.Cq[i], p.Nq[i] / q[PI_P_R_NC]), p.Pq[i] / q[PI_P_R_PC]);
            p.exu[i] = M::mask(p.ac[i], M::max(z0, p.Cq[i] - mn2) * q[PI_P_k_mtb]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_biomass_kernel! */
            p.Bm[i] = p.Bm[i] + dt * p.BS[i];
            p.Cq[i] = p.Cq[i] - dt * (p.BS[i] + p.exu[i]);
            p.Nq[i] = p.Nq[i] - dt * p.BS[i] * q[PI_P_R_NC];
            p.Pq[i] = p.Pq[i] - dt * p.BS[i] * q[PI_P_R_PC];
            p.Chl[i] = p.Chl[i] + dt * p.BS[i] * q[PI_P_R_NC] * p.rho[i];
            p.age[i] = p.age[i] + M::mask(p.ac[i], dt / L(3600));
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* update_cellsize_kernel! */
            p.Sz[i] = p.Bm[i] / (q[PI_P_Cquota] * q[PI_P_Nsuper]);
        /* probabilities (division_death_probability.jl:176-239) */
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_graz_quadratic_kernel! */
            p.graz[i] = M::mask(p.ac[i], tpop[i] * q[PI_P_grz_P]);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_mort_kernel! */
            p.mort[i] = M::mask(p.ac[i], q[PI_P_mort_P] * shape_dec(p.Sz[i], q[PI_P_mort_reg], L(1.0e-5f), 4));
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_dvid_kernel! with calc_division (:178-183) */
            FT P_ = q[PI_P_dvid_P], Sz = p.Sz[i], v;
            FT hour = M::rem(t, L(86400)) / L(3600);
            if (dvid_type[s] == 1) v = P_ * shape_inc(Sz, q[PI_P_dvid_reg], L(2.0e-3f), 2);
            else if (dvid_type[s] == 2) v = P_ * shape_inc(hour, q[PI_P_dvid_reg], L(1.0e-5f), 4);
            else v = P_ * shape_inc(Sz, q[PI_P_dvid_reg], L(1.0e-3f), 4) * shape_inc(hour, q[PI_P_dvid_reg2], L(1.0e-5f), 4);
            p.dvid[i] = M::mask(p.ac[i], v);
        }
        /* get_probability! (division_death_probability.jl:266-278) */
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            float a, b, c;
            pi_uniform3(seed, iteration, s, PI_KIND_DECISION, i, &a, &b, &c);
            bool ac = p.ac[i];
            p.graz[i] = (M::isless((FT)a, p.graz[i] * dt) && ac) ? L(1) : L(0);
            p.mort[i] = (M::isless((FT)b, p.mort[i] * dt) && ac) ? L(1) : L(0);
            p.dvid[i] = (M::isless((FT)c, p.dvid[i] * dt) && ac) ? L(1) : L(0);
        }
    }
    /* calc_consume_kernel! (QuotaMode/consume_loss.jl:2-14) */
    void calc_consume(const Plank& p, FT dt) {
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            int64_t c = cell_of(p, i);
            bool a = p.ac[i];
            plk[PI_DIC][c] += M::mask(a, (p.resp[i] - p.PS[i]) * dt);
            plk[PI_DOC][c] += M::mask(a, (p.exu[i] - p.VDOC[i]) * dt);
            plk[PI_NH4][c] += M::mask(a, -p.VNH4[i] * dt);
            plk[PI_NO3][c] += M::mask(a, -p.VNO3[i] * dt);
            plk[PI_PO4][c] += M::mask(a, -p.VPO4[i] * dt);
        }
    }

    /* ================================================ CarbonMode ======= */
    /* tempFunc_PS / tempFunc (CarbonMode/growth_kernels.jl:1-17): regT with k = 0.1 */
    static FT c_tempFunc_PS(FT T, const std::array<FT, PI_NPARAMS>& q) {
        FT x = T - q[PI_C_Topt], xmax = q[PI_C_Tmax] - q[PI_C_Topt];
        FT regT = shape_dec(x, xmax, L(1.0e-1f), 4);
        FT k = M::exp(-q[PI_C_Ea] / (L(8.3145f) * (T + L(273.15f)))) * regT;
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_C_Ea] / (L(8.3145f) * (q[PI_C_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }
    static FT c_tempFunc(FT T, const std::array<FT, PI_NPARAMS>& q) {
        FT k = M::exp(-q[PI_C_Ea] / (L(8.3145f) * (T + L(273.15f))));
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_C_Ea] / (L(8.3145f) * (q[PI_C_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }
    /* gamma_alloc (CarbonMode/growth_kernels.jl:22-26) */
    static FT gamma_alloc(FT Bm, FT Bd, const std::array<FT, PI_NPARAMS>& q) {
        FT g = M::mask(M::isless(L(0), Bd), Bd / M::max(L(1.0e-30f), Bm));
        return M::mask(q[PI_C_thermal] == L(1), M::min(L(1), g));
    }
    /* calc_loss_kernel! (CarbonMode/consume_loss.jl:11-21): DOC and POC only */
    void c_calc_loss(const Plank& p, const V& loss, FT fC) {
        for (int64_t i = 0; i < cap; ++i) {
            int64_t c = cell_of(p, i);
            bool a = p.ac[i];
            plk[PI_DOC][c] += M::mask(a, p.Bm[i] * fC) * loss[i];
            plk[PI_POC][c] += M::mask(a, p.Bm[i] * (L(1) - fC)) * loss[i];
        }
    }
    /* plankton_update! (CarbonMode/plankton_update.jl:1-32) */
    pi_status carbon_update(int s, FT dt) {
        Plank& p = sp[s];
        const auto& q = P[s];
        diags_proc(diagSp[s][1], p.graz, p);
        c_calc_loss(p, p.graz, q[PI_C_grazFracC]);
        inactivate(p, p.graz);
        diags_proc(diagSp[s][2], p.mort, p);
        c_calc_loss(p, p.mort, q[PI_C_mortFracC]);
        inactivate(p, p.mort);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.dvid[i] = M::mask(p.ac[i], p.dvid[i]);
        diags_proc(diagSp[s][3], p.dvid, p);
        FT dvidnum = 0;
        for (int64_t i = 0; i < cap; ++i) dvidnum += M::mask(p.ac[i], p.dvid[i]);
        std::vector<int64_t> free_ind;
        for (int64_t i = 0; i < cap; ++i) if (!p.ac[i]) free_ind.push_back(i + 1);
        if (dvidnum > (FT)free_ind.size()) {
            char buf[160];
            snprintf(buf, sizeof buf, "number of individual exceeds the capacity at timestep %g days",
                     (double)t / 86400.0);
            failed = true;
            return fail(PI_ERR_CAPACITY, buf);
        }
        c_divide(p, free_ind);
        c_growth(s, dt);
        for (int64_t i = 0; i < cap; ++i)    /* calc_consume_kernel! (consume_loss.jl:2-9): DIC only */
            plk[PI_DIC][cell_of(p, i)] += M::mask(p.ac[i], (p.resp[i] - p.PS[i]) * dt);
        for (int64_t i = 0; i < cap; ++i) diagSp[s][0][cell_of(p, i)] += (FT)((p.ac[i] && p.ac[i]) ? 1 : 0);
        return PI_OK;
    }
    /* divide! (CarbonMode/division_death.jl:28-70): the daughter copies x, y, z, Sz,
     * Bm, Chl, gen, ac, dvid; Bd and age stay those of the slot */
    void c_divide(Plank& p, const std::vector<int64_t>& free_ind) {
        FT acc = 0;
        for (int64_t i = 0; i < cap; ++i) { acc = acc + p.dvid[i]; idc[i] = acc; }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) idc_int[i] = (int64_t)idc[i];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)
            if (p.dvid[i] == L(1)) p.idx[i] = free_ind[(size_t)(idc_int[i] - 1)];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {
            if (p.dvid[i] == L(1) && p.idx[i] != 0) {
                int64_t d = p.idx[i] - 1;
                p.x[d] = p.x[i]; p.y[d] = p.y[i]; p.z[d] = p.z[i];
                p.Sz[d] = p.Sz[i]; p.Bm[d] = p.Bm[i]; p.Chl[d] = p.Chl[i];
                p.gen[d] = p.gen[i]; p.ac[d] = p.ac[i]; p.dvid[d] = p.dvid[i];
            }
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* divide_to_half_kernel! */
            FT h = (L(2) - p.dvid[i]) / L(2);
            p.Sz[i] = p.Sz[i] * h; p.Bm[i] = p.Bm[i] * h; p.Cq[i] = p.Cq[i] * h; p.Chl[i] = p.Chl[i] * h;
            p.gen[i] = p.gen[i] + p.dvid[i];
            p.age[i] = p.age[i] * (L(1) - p.dvid[i]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.idx[i] = 0;
    }
    /* plankton_growth! (CarbonMode/plankton_growth.jl:2-29, growth_kernels.jl:28-120,
     * division_death_probability.jl:176-278) */
    void c_growth(int s, FT dt) {
        Plank& p = sp[s];
        const auto& q = P[s];
        const FT e30 = L(1.0e-30f), z0 = L(0), one = L(1);
        const bool thermal_on = q[PI_C_thermal] == one;
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_PS_kernel! with calc_photosynthesis */
            FT aI = tpar[i] * q[PI_C_alpha] * q[PI_C_Phi];
            FT T = tT[i];
            FT PCm = q[PI_C_PCmax] * (c_tempFunc(T, q) * q[PI_C_thermal] + c_tempFunc_PS(T, q) * (one - q[PI_C_thermal]));
            FT light = one - M::exp(-aI / M::max(e30, PCm) * p.Chl[i] / M::max(e30, p.Bm[i]));
            FT PS = PCm * M::max(z0, p.Bm[i] - p.Cq[i]) * (light * (one - q[PI_C_is_bact]) + q[PI_C_is_bact]);
            p.PS[i] = M::mask(p.ac[i], PS);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_BS_kernel! */
            p.BS[i] = p.PS[i] * (one - gamma_alloc(p.Bm[i], p.Cq[i], q));
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_repair_kernel! */
            p.RP[i] = p.PS[i] * gamma_alloc(p.Bm[i], p.Cq[i], q);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_thermal_damage_kernel! */
            FT T = tT[i], Bm = p.Bm[i], Bd = p.Cq[i];
            FT td = (T - q[PI_C_Topt]) * q[PI_C_f_T2B] * (Bm - Bd);
            td = M::mask(M::isless(q[PI_C_Topt], T), td);
            td = M::mask(M::isless(Bd, Bm), td);
            td = M::mask(thermal_on, td);
            p.TD[i] = M::max(z0, M::min(td, (Bm - Bd) / dt));
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_respir_kernel! */
            p.resp[i] = M::mask(p.ac[i], q[PI_C_respir] * p.Bm[i] * c_tempFunc(tT[i], q));
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_quotas_kernel! */
            p.Bm[i] = p.Bm[i] + (p.BS[i] - p.resp[i]) * dt;
            p.Cq[i] = p.Cq[i] + (p.TD[i] - p.RP[i]) * dt;
            p.age[i] = p.age[i] + M::mask(p.ac[i], dt / L(3600));
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_cellsize_kernel! */
            p.Sz[i] = p.Bm[i] / (q[PI_C_Cquota] * q[PI_C_Nsuper]);
            p.Chl[i] = p.Bm[i] * q[PI_C_Chl2C];
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_graz_quadratic_kernel! */
            p.graz[i] = M::mask(p.ac[i], tpop[i] * q[PI_C_grz_P]);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_dvid_kernel! */
            FT P_ = q[PI_C_dvid_P], Sz = p.Sz[i], v;
            FT hour = M::rem(t, L(86400)) / L(3600);
            if (dvid_type[s] == 1) v = P_ * shape_inc(Sz, q[PI_C_dvid_reg], L(2.0e-3f), 2);
            else if (dvid_type[s] == 2) v = P_ * shape_inc(hour, q[PI_C_dvid_reg], L(1.0e-5f), 4);
            else v = P_ * shape_inc(Sz, q[PI_C_dvid_reg], L(1.0e-3f), 4) * shape_inc(hour, q[PI_C_dvid_reg2], L(1.0e-5f), 4);
            p.dvid[i] = M::mask(p.ac[i], v);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_thermal_mort_kernel! / calc_mort_kernel! */
            FT v = thermal_on
                ? q[PI_C_mort_P] * shape_inc(p.Cq[i] / M::max(e30, p.Bm[i]), L(0.99f), L(1.0e-6f), 4)
                : q[PI_C_mort_P] * shape_dec(p.Sz[i], q[PI_C_mort_reg], L(1.0e-5f), 4);
            p.mort[i] = M::mask(p.ac[i], v);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* get_probability! */
            float a, b, c;
            pi_uniform3(seed, iteration, s, PI_KIND_DECISION, i, &a, &b, &c);
            bool ac = p.ac[i];
            p.graz[i] = (M::isless((FT)a, p.graz[i] * dt) && ac) ? L(1) : L(0);
            p.mort[i] = (M::isless((FT)b, p.mort[i] * dt) && ac) ? L(1) : L(0);
            p.dvid[i] = (M::isless((FT)c, p.dvid[i] * dt) && ac) ? L(1) : L(0);
        }
    }

    /* ========================================== MacroMolecularMode ===== */
    using Q = std::array<FT, PI_NPARAMS>;
    /* total_C_biomass / total_N_biomass / total_P_biomass
     * (MacroMolecularMode/plankton_generation.jl:84-95) */
    static FT m_totC(FT PRO, FT DNA, FT RNA, FT CH, FT Chl) {
        return PRO + DNA + RNA + CH + Chl / L(893.49f) * L(55.0f);
    }
    static FT m_totN(FT PRO, FT DNA, FT RNA, FT NST, FT Chl, const Q& q) {
        return PRO * q[PI_M_R_NC_PRO] + DNA * q[PI_M_R_NC_DNA] + RNA * q[PI_M_R_NC_RNA] + NST + Chl / L(893.49f) * L(4.0f);
    }
    static FT m_totP(FT DNA, FT RNA, FT PST, const Q& q) {
        return DNA * q[PI_M_R_PC_DNA] + RNA * q[PI_M_R_PC_RNA] + PST;
    }
    /* tempFunc_PS / tempFunc (MacroMolecularMode/growth_kernels.jl:1-17) */
    static FT m_tempFunc_PS(FT T, const Q& q) {
        FT x = T - q[PI_M_Topt], xmax = q[PI_M_Tmax] - q[PI_M_Topt];
        FT regT = shape_dec(x, xmax, L(4.0e-2f), 4);
        FT k = M::exp(-q[PI_M_Ea] / (L(8.3145f) * (T + L(273.15f)))) * regT;
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_M_Ea] / (L(8.3145f) * (q[PI_M_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }
    static FT m_tempFunc(FT T, const Q& q) {
        FT k = M::exp(-q[PI_M_Ea] / (L(8.3145f) * (T + L(273.15f))));
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_M_Ea] / (L(8.3145f) * (q[PI_M_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }
    /* calc_loss_kernel! (MacroMolecularMode/consume_loss.jl:17-26) */
    void m_calc_loss(const Plank& p, const V& loss, FT fC, FT fN, FT fP, const Q& q) {
        const V &PRO = p.Bm, &DNA = p.Sz, &RNA = p.RNA, &CH = p.Cq, &NST = p.Nq, &PST = p.Pq;
        for (int64_t i = 0; i < cap; ++i) {
            int64_t c = cell_of(p, i);
            bool a = p.ac[i];
            FT C = m_totC(PRO[i], DNA[i], RNA[i], CH[i], p.Chl[i]);
            FT N = m_totN(PRO[i], DNA[i], RNA[i], NST[i], p.Chl[i], q);
            FT P_ = m_totP(DNA[i], RNA[i], PST[i], q);
            plk[PI_DOC][c] += M::mask(a, C * fC) * loss[i];
            plk[PI_DON][c] += M::mask(a, N * fN) * loss[i];
            plk[PI_DOP][c] += M::mask(a, P_ * fP) * loss[i];
            plk[PI_POC][c] += M::mask(a, C * (L(1) - fC)) * loss[i];
            plk[PI_PON][c] += M::mask(a, N * (L(1) - fN)) * loss[i];
            plk[PI_POP][c] += M::mask(a, P_ * (L(1) - fP)) * loss[i];
        }
    }
    /* plankton_update! (MacroMolecularMode/plankton_update.jl:1-33) */
    pi_status mm_update(int s, FT dt) {
        Plank& p = sp[s];
        const auto& q = P[s];
        diags_proc(diagSp[s][1], p.graz, p);
        m_calc_loss(p, p.graz, q[PI_M_grazFracC], q[PI_M_grazFracN], q[PI_M_grazFracP], q);
        inactivate(p, p.graz);
        diags_proc(diagSp[s][2], p.mort, p);
        m_calc_loss(p, p.mort, q[PI_M_mortFracC], q[PI_M_mortFracN], q[PI_M_mortFracP], q);
        inactivate(p, p.mort);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.dvid[i] = M::mask(p.ac[i], p.dvid[i]);
        diags_proc(diagSp[s][3], p.dvid, p);
        FT dvidnum = 0;
        for (int64_t i = 0; i < cap; ++i) dvidnum += M::mask(p.ac[i], p.dvid[i]);
        std::vector<int64_t> free_ind;
        for (int64_t i = 0; i < cap; ++i) if (!p.ac[i]) free_ind.push_back(i + 1);
        if (dvidnum > (FT)free_ind.size()) {
            char buf[160];
            snprintf(buf, sizeof buf, "number of individual exceeds the capacity at timestep %g days",
                     (double)t / 86400.0);
            failed = true;
            return fail(PI_ERR_CAPACITY, buf);
        }
        m_divide(p, free_ind);
        m_growth(s, dt);
        calc_consume(p, dt);   /* MacroMolecularMode/consume_loss.jl:2-14: as QuotaMode's */
        for (int64_t i = 0; i < cap; ++i) diagSp[s][0][cell_of(p, i)] += (FT)((p.ac[i] && p.ac[i]) ? 1 : 0);
        return PI_OK;
    }
    /* divide! (MacroMolecularMode/division_death.jl:28-78): the daughter copies x,
     * y, z, PRO, DNA, RNA, CH, NST, PST, Chl, gen, ac, dvid; age stays the slot's */
    void m_divide(Plank& p, const std::vector<int64_t>& free_ind) {
        V &PRO = p.Bm, &DNA = p.Sz, &RNA = p.RNA, &CH = p.Cq, &NST = p.Nq, &PST = p.Pq;
        FT acc = 0;
        for (int64_t i = 0; i < cap; ++i) { acc = acc + p.dvid[i]; idc[i] = acc; }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) idc_int[i] = (int64_t)idc[i];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)
            if (p.dvid[i] == L(1)) p.idx[i] = free_ind[(size_t)(idc_int[i] - 1)];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {                  /* copy_daughter_individuals_kernel! */
            if (p.dvid[i] == L(1) && p.idx[i] != 0) {
                int64_t d = p.idx[i] - 1;
                p.x[d] = p.x[i]; p.y[d] = p.y[i]; p.z[d] = p.z[i];
                PRO[d] = PRO[i]; DNA[d] = DNA[i]; RNA[d] = RNA[i]; CH[d] = CH[i]; NST[d] = NST[i]; PST[d] = PST[i];
                p.Chl[d] = p.Chl[i]; p.gen[d] = p.gen[i]; p.ac[d] = p.ac[i]; p.dvid[d] = p.dvid[i];
            }
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {                  /* divide_to_half_kernel! */
            FT h = (L(2) - p.dvid[i]) / L(2);
            PRO[i] = PRO[i] * h; DNA[i] = DNA[i] * h; RNA[i] = RNA[i] * h; CH[i] = CH[i] * h;
            NST[i] = NST[i] * h; PST[i] = PST[i] * h; p.Chl[i] = p.Chl[i] * h;
            p.gen[i] = p.gen[i] + p.dvid[i];
            p.age[i] = p.age[i] * (L(1) - p.dvid[i]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.idx[i] = 0;
    }
    /* plankton_growth! (MacroMolecularMode/plankton_growth.jl:2-28, growth_kernels.jl:19-200,
     * division_death_probability.jl:39-44,96-101,258-278), every kernel over every slot */
    void m_growth(int s, FT dt) {
        Plank& p = sp[s];
        const auto& q = P[s];
        V &PRO = p.Bm, &DNA = p.Sz, &RNA = p.RNA, &CH = p.Cq, &NST = p.Nq, &PST = p.Pq;
        const FT e30 = L(1.0e-30f), z0 = L(0), one = L(1);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_inorganic_uptake_kernel! */
            FT aI = tpar[i] * q[PI_M_alpha] * q[PI_M_Phi];                            /* calc_PS */
            FT PCm = q[PI_M_PCmax] * m_tempFunc_PS(tT[i], q);
            FT PS = PCm * (one - M::exp(-aI / M::max(e30, PCm) * p.Chl[i] / M::max(e30, PRO[i]))) * PRO[i];
            p.PS[i] = M::mask(p.ac[i], PS);
            FT N_tot = m_totN(PRO[i], DNA[i], RNA[i], NST[i], p.Chl[i], q);           /* calc_NP_uptake */
            FT P_tot = m_totP(DNA[i], RNA[i], PST[i], q);
            FT R_NST = NST[i] / M::max(e30, N_tot);
            FT R_PST = PST[i] / M::max(e30, P_tot);
            FT regQN = shape_dec(R_NST, q[PI_M_NSTmax], L(1.0e-4f), 4);
            FT regQP = shape_dec(R_PST, q[PI_M_PSTmax], L(1.0e-4f), 4);
            FT T = tT[i];
            bool a = p.ac[i];
            FT VNH4 = M::mask(a, q[PI_M_VNH4max] * regQN * tNH4[i] / M::max(e30, tNH4[i] + q[PI_M_KsatNH4]) * m_tempFunc(T, q) * PRO[i]);
            FT VNO3 = M::mask(a, q[PI_M_VNO3max] * regQN * tNO3[i] / M::max(e30, tNO3[i] + q[PI_M_KsatNO3]) * m_tempFunc(T, q) * PRO[i]);
            FT VPO4 = M::mask(a, q[PI_M_VPO4max] * regQP * tPO4[i] / M::max(e30, tPO4[i] + q[PI_M_KsatPO4]) * m_tempFunc(T, q) * PRO[i]);
            p.VNH4[i] = M::min(VNH4, tNH4[i] / dt / M::max(one, tpop[i]));
            p.VNO3[i] = M::min(VNO3, tNO3[i] / dt / M::max(one, tpop[i]));
            p.VPO4[i] = M::min(VPO4, tPO4[i] / dt / M::max(one, tpop[i]));
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_quotas_1_kernel! */
            CH[i] = CH[i] + p.PS[i] * dt;
            PST[i] = PST[i] + p.VPO4[i] * dt;
            NST[i] = NST[i] + (p.VNH4[i] + p.VNO3[i]) * dt;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_organic_uptake_kernel! (calc_DOC_uptake) */
            FT C_tot = m_totC(PRO[i], DNA[i], RNA[i], CH[i], p.Chl[i]);
            FT R_CH = CH[i] / M::max(e30, C_tot);
            FT regQ = shape_dec(R_CH, q[PI_M_CHmax], L(1.0e-4f), 4);
            FT VN = q[PI_M_VDOCmax] * regQ * tDOC[i] / M::max(e30, tDOC[i] + q[PI_M_KsatDOC]) * m_tempFunc(tT[i], q) * PRO[i];
            p.VDOC[i] = M::mask(p.ac[i], M::min(VN, tDOC[i] / dt / M::max(one, tpop[i])));
            p.VDOC[i] = M::mask(M::isless(L(0.05f), p.PS[i] / (p.VDOC[i] + p.PS[i])), p.VDOC[i]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_ρChl_kernel! */
            FT r = p.PS[i] / M::max(e30, PRO[i]) /
                   M::max(e30, tpar[i] * q[PI_M_alpha] * q[PI_M_Phi] * p.Chl[i] / M::max(e30, PRO[i]));
            r = M::mask(M::isless(L(0.1f), tpar[i]), r);
            p.rho[i] = M::mask(p.ac[i], r);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_respir_kernel! */
            p.resp[i] = M::mask(p.ac[i], q[PI_M_respir] * PRO[i] * m_tempFunc(tT[i], q));
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_quotas_2_kernel! */
            CH[i] = CH[i] + (p.VDOC[i] - p.resp[i]) * dt;
            PRO[i] = PRO[i] - M::max(z0, (z0 - CH[i]));
            NST[i] = NST[i] + M::max(z0, (z0 - CH[i])) * q[PI_M_R_NC_PRO];
            CH[i] = CH[i] + M::max(z0, (z0 - CH[i]));
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_BS_kernel! */
            const FT Ns = q[PI_M_Nsuper];
            FT limit_PRO = M::min(CH[i] / (CH[i] + q[PI_M_k_sat_pro] * Ns),
                                  NST[i] / (NST[i] + q[PI_M_k_sat_pro] * Ns * q[PI_M_R_NC_PRO]));
            FT limit_DNA = M::min(M::min(CH[i] / (CH[i] + q[PI_M_k_sat_dna] * Ns),
                                         NST[i] / (NST[i] + q[PI_M_k_sat_dna] * Ns * q[PI_M_R_NC_DNA])),
                                  PST[i] / (PST[i] + q[PI_M_k_sat_dna] * Ns * q[PI_M_R_PC_DNA]));
            FT limit_RNA = M::min(M::min(CH[i] / (CH[i] + q[PI_M_k_sat_rna] * Ns),
                                         NST[i] / (NST[i] + q[PI_M_k_sat_rna] * Ns * q[PI_M_R_NC_RNA])),
                                  PST[i] / (PST[i] + q[PI_M_k_sat_rna] * Ns * q[PI_M_R_PC_RNA]));
            FT tf = m_tempFunc(tT[i], q);
            p.S_PRO[i] = q[PI_M_k_pro] * RNA[i] * limit_PRO * tf;
            p.S_DNA[i] = M::mask(M::isless(DNA[i] / (q[PI_M_C_DNA] * Ns), L(2)), q[PI_M_k_dna] * PRO[i] * limit_DNA * tf);
            p.S_RNA[i] = q[PI_M_k_rna] * PRO[i] * limit_RNA * tf;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_biomass_kernel! */
            PRO[i] = PRO[i] + dt * p.S_PRO[i];
            DNA[i] = DNA[i] + dt * p.S_DNA[i];
            RNA[i] = RNA[i] + dt * p.S_RNA[i];
            CH[i] = CH[i] - dt * (p.S_PRO[i] + p.S_DNA[i] + p.S_RNA[i] + p.S_PRO[i] * p.rho[i]);
            NST[i] = NST[i] - dt * (p.S_PRO[i] * q[PI_M_R_NC_PRO] + p.S_DNA[i] * q[PI_M_R_NC_DNA] +
                                    p.S_RNA[i] * q[PI_M_R_NC_RNA] + p.S_PRO[i] * p.rho[i] * L(4.0f) / L(55.0f));
            PST[i] = PST[i] - dt * (p.S_DNA[i] * q[PI_M_R_PC_DNA] + p.S_RNA[i] * q[PI_M_R_PC_RNA]);
            p.Chl[i] = p.Chl[i] + dt * p.S_PRO[i] * p.rho[i] * L(893.49f) / L(55.0f);
            p.age[i] = p.age[i] + M::mask(p.ac[i], dt / L(3600));
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_exudation_kernel! */
            FT tot_C = m_totC(PRO[i], DNA[i], RNA[i], CH[i], p.Chl[i]);
            p.exu[i] = M::max(z0, CH[i] - q[PI_M_CHmax] * tot_C);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* update_CH_kernel! */
            CH[i] = CH[i] - p.exu[i];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_graz_quadratic_kernel! */
            p.graz[i] = M::mask(p.ac[i], tpop[i] * q[PI_M_grz_P]);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_MM_mort_kernel! */
            p.mort[i] = M::mask(p.ac[i], q[PI_M_mort_P] * shape_inc(p.age[i], q[PI_M_mort_reg], L(1.0e-4f), 4));
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_MM_dvid_kernel! */
            p.dvid[i] = M::mask(M::isless(L(2), DNA[i] / (q[PI_M_C_DNA] * q[PI_M_Nsuper])), q[PI_M_dvid_P]);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* get_probability! */
            float a, b, c;
            pi_uniform3(seed, iteration, s, PI_KIND_DECISION, i, &a, &b, &c);
            bool ac = p.ac[i];
            p.graz[i] = (M::isless((FT)a, p.graz[i] * dt) && ac) ? L(1) : L(0);
            p.mort[i] = (M::isless((FT)b, p.mort[i] * dt) && ac) ? L(1) : L(0);
            p.dvid[i] = (M::isless((FT)c, p.dvid[i] * dt) && ac) ? L(1) : L(0);
        }
    }

    /* ============================================ IronEnergyMode ======= */
    /* tempFunc_CF / tempFunc (IronEnergyMode/growth_kernels.jl:1-17) */
    static FT i_tempFunc_CF(FT T, const Q& q) {
        FT x = T - q[PI_I_Topt], xmax = q[PI_I_Tmax] - q[PI_I_Topt];
        FT regT = shape_dec(x, xmax, L(4.0e-2f), 4);
        FT k = M::exp(-q[PI_I_Ea] / (L(8.3145f) * (T + L(273.15f)))) * regT;
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_I_Ea] / (L(8.3145f) * (q[PI_I_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }
    static FT i_tempFunc(FT T, const Q& q) {
        FT k = M::exp(-q[PI_I_Ea] / (L(8.3145f) * (T + L(273.15f))));
        k = M::max(L(0), k);
        FT OGT = M::exp(-q[PI_I_Ea] / (L(8.3145f) * (q[PI_I_Topt] + L(273.15f))));
        return M::min(L(1), k / OGT);
    }
    /* calc_loss_kernel! (IronEnergyMode/consume_loss.jl:18-33) */
    void i_calc_loss(const Plank& p, const V& loss, FT fC, FT fN, FT fP, FT fFe, FT R_NC, FT R_PC) {
        for (int64_t i = 0; i < cap; ++i) {
            int64_t c = cell_of(p, i);
            bool a = p.ac[i];
            FT Fe = p.qFe[i] + p.qFePS[i] + p.qFeNF[i] + p.qFeNR[i];
            plk[PI_NH4][c] += M::mask(a, p.Nq[i]) * loss[i];
            plk[PI_NO3][c] += M::mask(a, p.qNO3[i]) * loss[i];
            plk[PI_DOC][c] += M::mask(a, (p.Bm[i] + p.Cq[i]) * fC) * loss[i];
            plk[PI_DON][c] += M::mask(a, p.Bm[i] * R_NC * fN) * loss[i];
            plk[PI_DOP][c] += M::mask(a, (p.Bm[i] * R_PC + p.Pq[i]) * fP) * loss[i];
            plk[PI_DFe][c] += M::mask(a, Fe * fFe) * loss[i];
            plk[PI_POC][c] += M::mask(a, (p.Bm[i] + p.Cq[i]) * (L(1) - fC)) * loss[i];
            plk[PI_PON][c] += M::mask(a, p.Bm[i] * R_NC * (L(1) - fN)) * loss[i];
            plk[PI_POP][c] += M::mask(a, (p.Bm[i] * R_PC + p.Pq[i]) * (L(1) - fP)) * loss[i];
            plk[PI_PFe_bio][c] += M::mask(a, Fe * (L(1) - fFe)) * loss[i];
        }
    }
    /* plankton_update! (IronEnergyMode/plankton_update.jl:1-33) */
    pi_status iron_update(int s, FT dt) {
        Plank& p = sp[s];
        const auto& q = P[s];
        diags_proc(diagSp[s][1], p.graz, p);
        i_calc_loss(p, p.graz, q[PI_I_grazFracC], q[PI_I_grazFracN], q[PI_I_grazFracP], q[PI_I_grazFracFe],
                    q[PI_I_R_NC], q[PI_I_R_PC]);
        inactivate(p, p.graz);
        diags_proc(diagSp[s][2], p.mort, p);
        i_calc_loss(p, p.mort, q[PI_I_mortFracC], q[PI_I_mortFracN], q[PI_I_mortFracP], q[PI_I_mortFracFe],
                    q[PI_I_R_NC], q[PI_I_R_PC]);
        inactivate(p, p.mort);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.dvid[i] = M::mask(p.ac[i], p.dvid[i]);
        diags_proc(diagSp[s][3], p.dvid, p);
        FT dvidnum = 0;
        for (int64_t i = 0; i < cap; ++i) dvidnum += M::mask(p.ac[i], p.dvid[i]);
        std::vector<int64_t> free_ind;
        for (int64_t i = 0; i < cap; ++i) if (!p.ac[i]) free_ind.push_back(i + 1);
        if (dvidnum > (FT)free_ind.size()) {
            char buf[160];
            snprintf(buf, sizeof buf, "number of individual exceeds the capacity at timestep %g days",
                     (double)t / 86400.0);
            failed = true;
            return fail(PI_ERR_CAPACITY, buf);
        }
        i_divide(p, free_ind);
        i_growth(s, dt);
        for (int64_t i = 0; i < cap; ++i) {   /* calc_consume_kernel! (consume_loss.jl:2-16) */
            int64_t c = cell_of(p, i);
            bool a = p.ac[i];
            plk[PI_DIC][c] += M::mask(a, (p.RS[i] - p.PS[i]) * dt);
            plk[PI_NH4][c] += M::mask(a, -p.VNH4[i] * dt);
            plk[PI_NO3][c] += M::mask(a, -p.VNO3[i] * dt);
            plk[PI_PO4][c] += M::mask(a, -p.VPO4[i] * dt);
            plk[PI_DFe][c] += M::mask(a, -p.VFe[i] * dt);
        }
        for (int64_t i = 0; i < cap; ++i) diagSp[s][0][cell_of(p, i)] += (FT)((p.ac[i] && p.ac[i]) ? 1 : 0);
        return PI_OK;
    }
    /* divide! (IronEnergyMode/division_death.jl:38-110): the daughter copies every
     * pool, exEn, Chl, gen, ac, dvid and tdark; age and ptc stay the slot's; both
     * halves get the halved pools and exEn (not tdark) */
    void i_divide(Plank& p, const std::vector<int64_t>& free_ind) {
        FT acc = 0;
        for (int64_t i = 0; i < cap; ++i) { acc = acc + p.dvid[i]; idc[i] = acc; }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) idc_int[i] = (int64_t)idc[i];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)
            if (p.dvid[i] == L(1)) p.idx[i] = free_ind[(size_t)(idc_int[i] - 1)];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {                  /* copy_daughter_individuals_kernel! */
            if (p.dvid[i] == L(1) && p.idx[i] != 0) {
                int64_t d = p.idx[i] - 1;
                p.x[d] = p.x[i]; p.y[d] = p.y[i]; p.z[d] = p.z[i];
                p.Sz[d] = p.Sz[i]; p.Bm[d] = p.Bm[i]; p.Cq[d] = p.Cq[i];
                p.qNO3[d] = p.qNO3[i]; p.Nq[d] = p.Nq[i]; p.Pq[d] = p.Pq[i];
                p.qFe[d] = p.qFe[i]; p.qFePS[d] = p.qFePS[i]; p.qFeNR[d] = p.qFeNR[i]; p.qFeNF[d] = p.qFeNF[i];
                p.exEn[d] = p.exEn[i]; p.Chl[d] = p.Chl[i]; p.gen[d] = p.gen[i]; p.ac[d] = p.ac[i];
                p.dvid[d] = p.dvid[i]; p.tdark[d] = p.tdark[i];
            }
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {                  /* divide_to_half_kernel! */
            FT h = (L(2) - p.dvid[i]) / L(2);
            p.Sz[i] = p.Sz[i] * h; p.Bm[i] = p.Bm[i] * h; p.Cq[i] = p.Cq[i] * h;
            p.qNO3[i] = p.qNO3[i] * h; p.Nq[i] = p.Nq[i] * h; p.Pq[i] = p.Pq[i] * h;
            p.qFe[i] = p.qFe[i] * h; p.qFePS[i] = p.qFePS[i] * h; p.qFeNR[i] = p.qFeNR[i] * h;
            p.qFeNF[i] = p.qFeNF[i] * h; p.exEn[i] = p.exEn[i] * h; p.Chl[i] = p.Chl[i] * h;
            p.gen[i] = p.gen[i] + p.dvid[i];
            p.age[i] = p.age[i] * (L(1) - p.dvid[i]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) p.idx[i] = 0;
    }
    /* plankton_growth! (IronEnergyMode/plankton_growth.jl:2-30, growth_kernels.jl:19-366,
     * division_death_probability.jl:26-95,266-278), every kernel over every slot.
     * CH = Cq, qNH4 = Nq, qP = Pq. */
    void i_growth(int s, FT dt) {
        Plank& p = sp[s];
        const auto& q = P[s];
        V &CH = p.Cq, &qNH4 = p.Nq, &qP = p.Pq;
        const FT e30 = L(1.0e-30f), z0 = L(0), one = L(1);
        const FT is_nr = q[PI_I_is_nr], is_nf = q[PI_I_is_croc] + q[PI_I_is_tric];
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_trs_uptake_kernel! */
            bool a = p.ac[i];
            FT T = tT[i], Bm = p.Bm[i], pp = M::max(one, tpop[i]);
            FT Qn = qNH4[i] / M::max(e30, Bm + CH[i]);                               /* calc_NH4_uptake */
            FT regQN = shape_dec(Qn, q[PI_I_qNH4max], L(1.0e-4f), 2);
            FT V1 = M::mask(a, q[PI_I_VNH4max] * regQN * tNH4[i] / M::max(e30, tNH4[i] + q[PI_I_KsatNH4]) * i_tempFunc(T, q) * Bm);
            p.VNH4[i] = M::min(V1, tNH4[i] / dt / pp) * is_nr;
            FT Qn3 = p.qNO3[i] / M::max(e30, Bm + CH[i]);                            /* calc_NO3_uptake */
            FT regQN3 = shape_dec(Qn3, q[PI_I_qNO3max], L(1.0e-4f), 2);
            FT V2 = M::mask(a, q[PI_I_VNO3max] * regQN3 * tNO3[i] / M::max(e30, tNO3[i] + q[PI_I_KsatNO3]) * i_tempFunc(T, q) * Bm);
            p.VNO3[i] = M::min(V2, tNO3[i] / dt / pp) * is_nr;
            FT Qp = (qP[i] + Bm * q[PI_I_R_PC]) / M::max(e30, Bm + CH[i]);            /* calc_P_uptake */
            FT regQP = shape_dec(Qp, q[PI_I_qPmax], L(1.0e-4f), 2);
            FT V3 = M::mask(a, q[PI_I_VPO4max] * regQP * tPO4[i] / M::max(e30, tPO4[i] + q[PI_I_KsatPO4]) * i_tempFunc(T, q) * Bm);
            p.VPO4[i] = M::min(V3, tPO4[i] / dt / pp);
            FT Qfe = (p.qFe[i] + p.qFePS[i] + p.qFeNF[i] + p.qFeNR[i]) / M::max(e30, Bm + CH[i]);   /* calc_Fe_uptake */
            FT regQFe = shape_dec(Qfe, q[PI_I_qFemax], L(1.0e-4f), 2);
            FT SA = q[PI_I_SA] * M::pow(p.Sz[i], L(2.0f) / L(3.0f));
            FT V4 = M::mask(a, q[PI_I_KSAFe] * SA * tDFe[i] * regQFe * q[PI_I_Nsuper] * i_tempFunc(T, q));
            p.VFe[i] = M::min(V4, tDFe[i] / dt / pp);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_state_1_kernel! */
            qP[i] = qP[i] + p.VPO4[i] * dt;
            p.qNO3[i] = p.qNO3[i] + p.VNO3[i] * dt;
            qNH4[i] = qNH4[i] + p.VNH4[i] * dt;
            p.qFe[i] = p.qFe[i] + p.VFe[i] * dt;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_PS_kernel! (calc_PS) */
            FT aI = tpar[i] * q[PI_I_alpha] * p.Chl[i] / M::max(e30, p.Bm[i]);
            FT Qfe_ps = p.qFePS[i] / M::max(e30, p.Bm[i] + CH[i]);
            FT Ksat = Qfe_ps / M::max(e30, Qfe_ps + q[PI_I_KfePS]);
            p.PS[i] = q[PI_I_PCmax] * (one - M::exp(-aI)) * p.Bm[i] * Ksat;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_respiration_kernel! (calc_respir) */
            bool a = p.ac[i];
            FT RS = M::mask(a, CH[i] * q[PI_I_k_rs] * i_tempFunc(tT[i], q));
            RS = M::min(RS, CH[i] / dt);
            p.RS[i] = RS;
            p.ERS[i] = M::mask(a, RS * q[PI_I_e_rs]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_carbon_fixation_kernel! (calc_CF) */
            bool a = p.ac[i];
            FT Qc = CH[i] / M::max(e30, p.Bm[i] + CH[i]);
            FT regQC = shape_dec(Qc, q[PI_I_CHmax], L(1.0e-4f), 2);
            p.CF[i] = M::mask(a, q[PI_I_k_cf] * regQC * i_tempFunc_CF(tT[i], q) * p.Bm[i]);
            p.ECF[i] = M::mask(a, p.CF[i] * q[PI_I_e_cf]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_NO3_reduction_kernel! (calc_NR) */
            bool a = p.ac[i];
            FT Qn = qNH4[i] / M::max(e30, p.Bm[i] + CH[i]);
            FT reg = shape_dec(Qn, q[PI_I_qNH4max], L(1.0e-4f), 2);
            FT Qfe_NR = p.qFeNR[i] / M::max(e30, p.Bm[i] + CH[i]);
            FT Ksat = Qfe_NR / M::max(e30, Qfe_NR + q[PI_I_KfeNR]);
            FT NR = M::mask(a, q[PI_I_k_nr] * reg * p.qNO3[i] * Ksat * i_tempFunc(tT[i], q));
            NR = M::min(NR, p.qNO3[i] / dt);
            FT ENR = M::mask(a, NR * q[PI_I_e_nr]);
            p.NR[i] = NR * is_nr;
            p.ENR[i] = ENR * is_nr;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_nitrogen_fixation_kernel! (calc_NF) */
            bool a = p.ac[i];
            FT Qn = qNH4[i] / M::max(e30, p.Bm[i] + CH[i]);
            FT reg = shape_dec(Qn, q[PI_I_qNH4max], L(1.0e-4f), 2);
            FT Qfe_NF = p.qFeNF[i] / M::max(e30, p.Bm[i] + CH[i]);
            FT Ksat = Qfe_NF / M::max(e30, Qfe_NF + q[PI_I_KfeNF]);
            FT NF = M::mask(a, q[PI_I_k_nf] * reg * Ksat * i_tempFunc(tT[i], q) * p.Bm[i]);
            FT ENF = M::mask(a, NF * q[PI_I_e_nf]);
            p.NF[i] = NF * is_nf;
            p.ENF[i] = ENF * is_nf;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* energy_allocation_kernel! (energy_alloc) */
            FT PS = p.PS[i], ERS = p.ERS[i];
            FT ECFt = M::min(PS, p.ECF[i]);
            FT ENFt = M::min(p.ENF[i], PS + ERS - ECFt) * is_nf;
            FT ENRt = M::min(p.ENR[i], PS + ERS - ECFt) * is_nr;
            FT ERSt = M::min(ERS, M::max(z0, ECFt + ENFt + ENRt - PS));
            p.exEn[i] = M::max(z0, PS - ECFt - ENFt - ENRt);
            p.ERS[i] = ERSt; p.ECF[i] = ECFt; p.ENF[i] = ENFt; p.ENR[i] = ENRt;
            p.RS[i] = p.ERS[i] / q[PI_I_e_rs];
            p.CF[i] = p.ECF[i] / q[PI_I_e_cf];
            p.NF[i] = p.ENF[i] / q[PI_I_e_nf];
            p.NR[i] = p.ENR[i] / q[PI_I_e_nr];
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_state_2_kernel! */
            CH[i] = CH[i] + (p.CF[i] - p.RS[i]) * dt;
            p.qNO3[i] = p.qNO3[i] + -p.NR[i] * dt;
            qNH4[i] = qNH4[i] + p.NR[i] * dt;
            qNH4[i] = qNH4[i] + p.NF[i] * dt;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_ρChl_kernel! */
            FT r = q[PI_I_Chl2N] / M::max(e30, tpar[i] * q[PI_I_alpha] * p.Chl[i] / p.Bm[i]);
            r = M::mask(M::isless(L(1.0e-1f), tpar[i]), r);
            p.rho[i] = M::mask(p.ac[i], r);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_BS_kernel! */
            p.BS[i] = M::mask(p.ac[i], M::min(M::min(CH[i], qNH4[i] / q[PI_I_R_NC]), qP[i] / q[PI_I_R_PC]) * q[PI_I_k_mtb]);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_iron_fluxes_kernel! (iron_alloc); Bool sums are Int */
            FT par = tpar[i], dpar = tdpar[i], qFe = p.qFe[i], qFePS = p.qFePS[i];
            FT qFeNR = p.qFeNR[i], qFeNF = p.qFeNF[i], tdark = p.tdark[i];
            int dark = M::isequal(L(0), par), light = M::isless(L(0), par);
            int rising = M::isless(L(0), dpar), falling = M::isless(dpar, L(0));
            int early = M::isless(tdark, q[PI_I_NF_clock]), late = M::isless(q[PI_I_NF_clock], tdark);
            FT f_ST2PS = M::mask(rising, q[PI_I_k_Fe_ST2PS] * qFe);
            FT f_PS2ST = q[PI_I_k_Fe_PS2ST] * qFePS * (FT)(falling + dark);
            FT f_ST2NR = M::mask(early, M::mask(dark, q[PI_I_k_Fe_ST2NR] * qFe)) * q[PI_I_is_nr];
            FT f_NR2ST = q[PI_I_k_Fe_NR2ST] * qFeNR * (FT)((dark && late) + light) * q[PI_I_is_nr];
            FT f_ST2NF_cr = M::mask(early, M::mask(dark, q[PI_I_k_Fe_ST2NF] * qFe));
            FT f_NF2ST_cr = q[PI_I_k_Fe_NF2ST] * qFeNF * (FT)((dark && late) + light);
            FT f_ST2NF_tr = M::mask(rising, q[PI_I_k_Fe_ST2NF] * qFe);
            FT f_NF2ST_tr = q[PI_I_k_Fe_NF2ST] * qFeNF * (FT)(falling + dark);
            FT f_ST2NF = f_ST2NF_cr * q[PI_I_is_croc] + f_ST2NF_tr * q[PI_I_is_tric];
            FT f_NF2ST = f_NF2ST_cr * q[PI_I_is_croc] + f_NF2ST_tr * q[PI_I_is_tric];
            f_ST2PS = M::min(f_ST2PS, qFe / dt);
            f_PS2ST = M::min(f_PS2ST, qFePS / dt);
            f_ST2NR = M::min(f_ST2NR, qFe / dt - f_ST2PS);
            f_NR2ST = M::min(f_NR2ST, qFeNR / dt);
            f_ST2NF = M::min(f_ST2NF, qFe / dt - f_ST2PS);
            f_NF2ST = M::min(f_NF2ST, qFeNF / dt);
            p.ST2PS[i] = f_ST2PS; p.PS2ST[i] = f_PS2ST; p.ST2NR[i] = f_ST2NR;
            p.NR2ST[i] = f_NR2ST; p.ST2NF[i] = f_ST2NF; p.NF2ST[i] = f_NF2ST;
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* update_biomass_kernel! */
            p.Bm[i] = p.Bm[i] + dt * p.BS[i];
            CH[i] = CH[i] - dt * p.BS[i];
            qNH4[i] = qNH4[i] - dt * p.BS[i] * q[PI_I_R_NC];
            qP[i] = qP[i] - dt * p.BS[i] * q[PI_I_R_PC];
            p.Chl[i] = p.Chl[i] + dt * p.BS[i] * q[PI_I_R_NC] * p.rho[i];
            p.age[i] = p.age[i] + M::mask(p.ac[i], dt / L(3600));
            p.qFe[i] = p.qFe[i] + dt * (p.PS2ST[i] - p.ST2PS[i] + p.NR2ST[i] - p.ST2NR[i] + p.NF2ST[i] - p.ST2NF[i]);
            p.qFePS[i] = p.qFePS[i] + dt * (p.ST2PS[i] - p.PS2ST[i]);
            p.qFeNR[i] = p.qFeNR[i] + dt * (p.ST2NR[i] - p.NR2ST[i]);
            p.qFeNF[i] = p.qFeNF[i] + dt * (p.ST2NF[i] - p.NF2ST[i]);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* update_cellsize_kernel! */
            p.Sz[i] = p.Bm[i] / (q[PI_I_Cquota] * q[PI_I_Nsuper]);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* update_tdark_kernel! */
            p.tdark[i] = p.tdark[i] * (one - (FT)M::isless(L(0), tpar[i])) + M::mask(M::isequal(L(0), tpar[i]), dt);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_graz_quadratic_kernel! */
            p.graz[i] = M::mask(p.ac[i], tpop[i] * q[PI_I_grz_P]);
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i)     /* calc_mort_kernel! */
            p.mort[i] = M::mask(p.ac[i], q[PI_I_mort_P] * shape_dec(p.Sz[i], q[PI_I_mort_reg], L(1.0e-5f), 4));
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* calc_dvid_kernel! */
            FT P_ = q[PI_I_dvid_P], Sz = p.Sz[i], v;
            FT hour = M::rem(t, L(86400)) / L(3600);
            if (dvid_type[s] == 1) v = P_ * shape_inc(Sz, q[PI_I_dvid_reg], L(2.0e-3f), 2);
            else if (dvid_type[s] == 2) v = P_ * shape_inc(hour, q[PI_I_dvid_reg], L(1.0e-5f), 4);
            else v = P_ * shape_inc(Sz, q[PI_I_dvid_reg], L(1.0e-3f), 4) * shape_inc(hour, q[PI_I_dvid_reg2], L(1.0e-5f), 4);
            p.dvid[i] = M::mask(p.ac[i], v);
        }
        OMP_FOR
        for (int64_t i = 0; i < cap; ++i) {   /* get_probability! */
            float a, b, c;
            pi_uniform3(seed, iteration, s, PI_KIND_DECISION, i, &a, &b, &c);
            bool ac = p.ac[i];
            p.graz[i] = (M::isless((FT)a, p.graz[i] * dt) && ac) ? L(1) : L(0);
            p.mort[i] = (M::isless((FT)b, p.mort[i] * dt) && ac) ? L(1) : L(0);
            p.dvid[i] = (M::isless((FT)c, p.dvid[i] * dt) && ac) ? L(1) : L(0);
        }
    }

    /* ===================== phytoplankton-particle interaction and release == */
    /* calc_distance (Abiotic/particle_interaction.jl:3-14) */
    FT calc_distance(int64_t xi1, int64_t yi1, int64_t zi1, FT x1, FT y1, FT z1,
                     int64_t xi2, int64_t yi2, int64_t zi2, FT x2, FT y2, FT z2) const {
        if (xi1 == xi2 && yi1 == yi2 && zi1 == zi2) {
            FT dx = (x1 - x2) * DxF((int)xi1, (int)yi1);
            FT dy = (y1 - y2) * DyF((int)xi1, (int)yi1);
            FT dz = (z1 - z2) * dzF_at((int)zi1);
            return (FT)std::sqrt(M::pow2(dx) + M::pow2(dy) + M::pow2(dz));
        }
        return L(1.0e20f);
    }
    /* particle_interaction! (:17-84): reset_interaction!, calc_interaction_kernel!
     * (every individual against every particle, later individuals overwriting the
     * candidate of their residue), merge_particle_kernel! (each particle in turn
     * joins the first candidate below max_ptc) */
    void particle_interaction(int s, int a) {
        Plank& pl = sp[s];
        Plank& pa = ab[a];
        const FT Rd = AP[a][PI_A_Rd], maxptc = P[s][PI_I_max_ptc];
        std::fill(intac.begin(), intac.end(), (int64_t)0);
        for (int64_t i = 0; i < cap; ++i) {
            if (!pl.ac[i]) continue;
            for (int64_t k = 0; k < cap; ++k) {
                if (!pa.ac[k]) continue;
                FT dist = calc_distance(pl.xi[i], pl.yi[i], pl.zi[i], pl.x[i], pl.y[i], pl.z[i],
                                        pa.xi[k], pa.yi[k], pa.zi[k], pa.x[k], pa.y[k], pa.z[k]);
                if (dist < Rd) {
                    const int64_t slot = ((i + 1) % mc) + 1;
                    intac[(size_t)(slot - 1) + (size_t)mc * k] = i + 1;
                }
            }
        }
        for (int64_t k = 0; k < cap; ++k) {
            if (!pa.ac[k]) continue;
            for (int j = 1; j <= mc; ++j) {
                const int64_t id = intac[(size_t)(j - 1) + (size_t)mc * k];
                if (id > 0 && pl.ptc[id - 1] < maxptc) {
                    pl.ptc[id - 1] += L(1);
                    pa.ac[k] = 0;
                    break;
                }
            }
        }
    }
    /* particle_release! (:86-145): get_release_probability! (rnd.x = this call's
     * draw), the capacity check, get_tind!, release_abiotic_particle_kernel! at
     * x + rnd.x, y + rnd.y, z + rnd.z with rnd.y, rnd.z still the last species'
     * draws of get_probability! (mortality and division); ptc of the other modes
     * is 0, so they release nothing */
    pi_status particle_release(int s, int a, int q, FT dt) {
        if (mode != PI_MODE_IRONENERGY) return PI_OK;
        Plank& pl = sp[s];
        Plank& pa = ab[a];
        const FT rP = AP[a][PI_A_release_P];
        V R(cap, (FT)0), rx(cap, (FT)0);
        FT releasenum = 0;
        for (int64_t i = 0; i < cap; ++i) {
            float u0, u1, u2;
            pi_uniform3(seed, iteration, PI_RELEASE_SPECIES + q, PI_KIND_RELEASE, i, &u0, &u1, &u2);
            rx[i] = (FT)u0;
            R[i] = (M::isless((FT)u0, rP * dt * pl.ptc[i]) && pl.ac[i]) ? L(1) : L(0);
        }
        for (int64_t i = 0; i < cap; ++i) releasenum += M::mask(pl.ac[i], R[i]);
        std::vector<int64_t> free_ind;
        for (int64_t k = 0; k < cap; ++k) if (!pa.ac[k]) free_ind.push_back(k + 1);
        if (releasenum > (FT)free_ind.size()) {
            char buf[160];
            snprintf(buf, sizeof buf, "number of abiotic particles exceeds the capacity at timestep %g days",
                     (double)t / 86400.0);
            failed = true;
            return fail(PI_ERR_CAPACITY, buf);
        }
        FT acc = 0;
        for (int64_t i = 0; i < cap; ++i) {
            acc = acc + R[i];
            if (R[i] == L(1)) {
                const int64_t k = free_ind[(size_t)((int64_t)acc - 1)] - 1;
                float b0, b1, b2;
                pi_uniform3(seed, iteration, nsp - 1, PI_KIND_DECISION, i, &b0, &b1, &b2);
                pa.ac[k] = 1;
                pa.x[k] = pl.x[i] + rx[i];
                pa.y[k] = pl.y[i] + (FT)b1;
                pa.z[k] = pl.z[i] + (FT)b2;
                pl.ptc[i] -= L(1);
            }
        }
        return PI_OK;
    }

    /* ================================================== tracers ========= */
    /* Difference operators (Biogeochemistry/Advection/operaters.jl:1-27) */
    inline FT dxp(const FT* c, int i, int j, int k) const { return M::mask(LM(i, j, k), M::mask(LM(i + 1, j, k), c[I(i + 1, j, k)] - c[I(i, j, k)])); }
    inline FT dyp(const FT* c, int i, int j, int k) const { return M::mask(LM(i, j, k), M::mask(LM(i, j + 1, k), c[I(i, j + 1, k)] - c[I(i, j, k)])); }
    inline FT dzp(const FT* c, int i, int j, int k) const { return M::mask(LM(i, j, k), M::mask(LM(i, j, k + 1), c[I(i, j, k)] - c[I(i, j, k + 1)])); }
    inline FT dx0(const FT* c, int i, int j, int k) const { return M::mask(LM(i - 1, j, k), M::mask(LM(i, j, k), c[I(i, j, k)] - c[I(i - 1, j, k)])); }
    inline FT dy0(const FT* c, int i, int j, int k) const { return M::mask(LM(i, j - 1, k), M::mask(LM(i, j, k), c[I(i, j, k)] - c[I(i, j - 1, k)])); }
    inline FT dz0(const FT* c, int i, int j, int k) const { return M::mask(LM(i, j, k - 1), M::mask(LM(i, j, k), c[I(i, j, k - 1)] - c[I(i, j, k)])); }
    inline FT dxm(const FT* c, int i, int j, int k) const { return M::mask(LM(i - 2, j, k), M::mask(LM(i - 1, j, k), c[I(i - 1, j, k)] - c[I(i - 2, j, k)])); }
    inline FT dym(const FT* c, int i, int j, int k) const { return M::mask(LM(i, j - 2, k), M::mask(LM(i, j - 1, k), c[I(i, j - 1, k)] - c[I(i, j - 2, k)])); }
    inline FT dzm(const FT* c, int i, int j, int k) const { return M::mask(LM(i, j, k - 2), M::mask(LM(i, j, k - 1), c[I(i, j, k - 2)] - c[I(i, j, k - 1)])); }

    /* DST3FL building blocks (Biogeochemistry/Advection/DST3FL.jl:1-73) */
    static FT d0(FT C) { return (L(2) - C) * (L(1) - C) / L(6); }
    static FT d1(FT C) { return (L(1) - C * C) / L(6); }
    static FT theta(FT dnum, FT dden) {
        const FT tmax = L(1.0e20f);
        return std::fabs(dden) * tmax <= std::fabs(dnum) ? std::copysign(tmax, dnum * dden) : dnum / dden;
    }
    static FT Psi(FT C, FT th) {
        return M::max(L(0), M::min(M::min(L(1), d0(C) + d1(C) * th), th * (L(1) - C) / (C + L(1.0e-20f))));
    }
    FT flux_x(const FT* c, const FT* u, int i, int j, int k, FT dt) const {
        FT C = std::fabs(u[I(i, j, k)] * dt / DxC(i, j));
        FT T = Ax(i, j, k) * u[I(i, j, k)];
        FT d = dx0(c, i, j, k);
        FT pp = Psi(C, theta(dxm(c, i, j, k), d));
        FT pm = Psi(C, theta(dxp(c, i, j, k), d));
        return L(0.5f) * (T + std::fabs(T)) * (c[I(i - 1, j, k)] + pp * d) +
               L(0.5f) * (T - std::fabs(T)) * (c[I(i, j, k)] - pm * d);
    }
    FT flux_y(const FT* c, const FT* v, int i, int j, int k, FT dt) const {
        FT C = std::fabs(v[I(i, j, k)] * dt / DyC(i, j));
        FT T = Ay(i, j, k) * v[I(i, j, k)];
        FT d = dy0(c, i, j, k);
        FT pp = Psi(C, theta(dym(c, i, j, k), d));
        FT pm = Psi(C, theta(dyp(c, i, j, k), d));
        return L(0.5f) * (T + std::fabs(T)) * (c[I(i, j - 1, k)] + pp * d) +
               L(0.5f) * (T - std::fabs(T)) * (c[I(i, j, k)] - pm * d);
    }
    FT flux_z(const FT* c, const FT* w, int i, int j, int k, FT dt) const {
        FT C = std::fabs(w[I(i, j, k)] * dt / dzC_at(k));
        FT T = Az(i, j) * w[I(i, j, k)];
        FT d = dz0(c, i, j, k);
        FT pp = Psi(C, theta(dzm(c, i, j, k), d));
        FT pm = Psi(C, theta(dzp(c, i, j, k), d));
        return L(0.5f) * (T + std::fabs(T)) * (c[I(i, j, k)] + pm * d) +
               L(0.5f) * (T - std::fabs(T)) * (c[I(i, j, k - 1)] - pp * d);
    }
    /* tracer_advection! (Biogeochemistry/Advection/multi_dim_adv.jl:111-135) */
    void tracer_advection(FT dt) {
        const V* vel = vel1;
        for (int q = 0; q < PI_NTRACERS; ++q) ttemp[q] = tr[q];
        for (int dim = 0; dim < 3; ++dim) {
            const FT* w = vel[dim].data();
            for (int q = 0; q < PI_NTRACERS; ++q)          /* calc_Gcs{x,y,z}! */
                OMP_FOR
                for (int k = Hz + 1; k <= Hz + Nz; ++k) for (int j = Hy + 1; j <= Hy + Ny; ++j)
                    for (int i = Hx + 1; i <= Hx + Nx; ++i)
                        Gcs[q][I(i, j, k)] = dim == 0 ? flux_x(ttemp[q].data(), w, i, j, k, dt)
                                           : dim == 1 ? flux_y(ttemp[q].data(), w, i, j, k, dt)
                                                      : flux_z(ttemp[q].data(), w, i, j, k, dt);
            fill_halo_Gcs(Gcs);
            for (int q = 0; q < PI_NTRACERS; ++q) {        /* multi_dim_{x,y,z}! */
                const FT* G = Gcs[q].data();
                const FT* c = tr[q].data();
                OMP_FOR
                for (int k = Hz + 1; k <= Hz + Nz; ++k) for (int j = Hy + 1; j <= Hy + Ny; ++j)
                    for (int i = Hx + 1; i <= Hx + Nx; ++i) {
                        FT dG, dT;
                        if (dim == 0) {
                            dG = dxp(G, i, j, k);
                            dT = M::mask(LM(i, j, k), M::mask(LM(i + 1, j, k), Ax(i + 1, j, k) * w[I(i + 1, j, k)] - Ax(i, j, k) * w[I(i, j, k)]));
                        } else if (dim == 1) {
                            dG = dyp(G, i, j, k);
                            dT = M::mask(LM(i, j, k), M::mask(LM(i, j + 1, k), Ay(i, j + 1, k) * w[I(i, j + 1, k)] - Ay(i, j, k) * w[I(i, j, k)]));
                        } else {
                            dG = dzp(G, i, j, k);
                            dT = M::mask(LM(i, j, k), M::mask(LM(i, j, k + 1), Az(i, j) * w[I(i, j, k)] - Az(i, j) * w[I(i, j, k + 1)]));
                        }
                        FT& tt = ttemp[q][I(i, j, k)];
                        tt = tt - dt / vol(i, j, k) * (dG - c[I(i, j, k)] * dT);
                    }
            }
            fill_halo_tracer(ttemp);
        }
        for (int q = 0; q < PI_NTRACERS; ++q)             /* calc_tracer_tendency! */
            OMP_FOR
            for (int64_t n = 0; n < ncell; ++n) Gcs[q][n] = ttemp[q][n] - tr[q][n];
    }
    /* tracer_diffusion! (Biogeochemistry/Advection/tracer_diffusion.jl:1-30) */
    void tracer_diffusion(FT dt) {
        const FT kx = B[PI_B_kappa_h], ky = B[PI_B_kappa_h], kz = B[PI_B_kappa_v];
        for (int q = 0; q < PI_NTRACERS; ++q) {
            const FT* c = tr[q].data();
            auto fx = [&](int i, int j, int k) { return kx * Ax(i, j, k) * (dx0(c, i, j, k) / DxF(i, j)); };
            auto fy = [&](int i, int j, int k) { return ky * Ay(i, j, k) * (dy0(c, i, j, k) / DyF(i, j)); };
            auto fz = [&](int i, int j, int k) { return kz * Az(i, j) * (dz0(c, i, j, k) / dzF_at(k)); };
            OMP_FOR
            for (int k = Hz + 1; k <= Hz + Nz; ++k) for (int j = Hy + 1; j <= Hy + Ny; ++j)
                for (int i = Hx + 1; i <= Hx + Nx; ++i) {
                    FT ddx = M::mask(LM(i, j, k), M::mask(LM(i + 1, j, k), fx(i + 1, j, k) - fx(i, j, k)));
                    FT ddy = M::mask(LM(i, j, k), M::mask(LM(i, j + 1, k), fy(i, j + 1, k) - fy(i, j, k)));
                    FT ddz = M::mask(LM(i, j, k), M::mask(LM(i, j, k + 1), fz(i, j, k) - fz(i, j, k + 1)));
                    FT lap = L(1) / vol(i, j, k) * (ddx + ddy + ddz);
                    FT& G = Gcs[q][I(i, j, k)];
                    G = G + lap * dt;
                }
        }
    }
    /* tracer_sinking! (Biogeochemistry/Advection/tracer_sinking.jl:1-43) */
    void tracer_sinking(FT dt) {
        const int names[3] = {PI_PFe_inorg, PI_Dust, PI_PFe_bio};
        for (int a = 0; a < 3; ++a) {
            int q = names[a];
            FT ws = a < 2 ? B[PI_B_w_sink_inorg] : B[PI_B_w_sink_org];
            OMP_FOR
            for (int64_t n = 0; n < ncell; ++n) flux_sink[n] = ws * tr[q][n];
            fill_top_flux_sink(flux_sink.data());
            fill_bottom_flux_sink(flux_sink.data());
            OMP_FOR
            for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
                int ii = i + Hx, jj = j + Hy, kk = k + Hz;
                FT& G = Gcs[q][I(ii, jj, kk)];
                G = G + (flux_sink[I(ii, jj, kk - 1)] - flux_sink[I(ii, jj, kk)]) / dzC_at(k) * dt;
            }
        }
    }
    /* tracer_forcing! (Biogeochemistry/tracer_forcings.jl:2-50), over padded arrays */
    void tracer_forcing(FT dt) {
        const FT z0 = L(0);
        for (int q = 0; q < PI_NTRACERS; ++q)
            OMP_FOR
            for (int64_t n = 0; n < ncell; ++n) ttemp[q][n] = M::max(z0, tr[q][n]);
        const FT kDOC = B[PI_B_kDOC], kPOC = B[PI_B_kPOC], kDON = B[PI_B_kDON], kPON = B[PI_B_kPON];
        const FT kDOP = B[PI_B_kDOP], kPOP = B[PI_B_kPOP], Nit = B[PI_B_Nit];
        const FT lPOC = B[PI_B_lambda_POC], lmin = B[PI_B_lambda_min], ldust = B[PI_B_lambda_dust];
        const FT DFeFrac = B[PI_B_DFeFrac], kdiss = B[PI_B_kdiss], lig = B[PI_B_ligand], lFe = B[PI_B_lambda_Fe];
        auto T = [&](int q, int64_t n) { return ttemp[q][n]; };
        OMP_FOR
        for (int64_t n = 0; n < ncell; ++n) {
            FT* F[PI_NTRACERS];
            for (int q = 0; q < PI_NTRACERS; ++q) F[q] = &Gcs[q][n];
            *F[PI_DIC] = *F[PI_DIC] + T(PI_DOC, n) * kDOC * dt;
            *F[PI_DOC] = *F[PI_DOC] - T(PI_DOC, n) * kDOC * dt + T(PI_POC, n) * kPOC * dt;
            *F[PI_POC] = *F[PI_POC] - T(PI_POC, n) * kPOC * dt;
            *F[PI_NH4] = *F[PI_NH4] + T(PI_DON, n) * kDON * dt - T(PI_NH4, n) * Nit * dt;
            *F[PI_NO3] = *F[PI_NO3] + T(PI_NH4, n) * Nit * dt;
            *F[PI_DON] = *F[PI_DON] - T(PI_DON, n) * kDON * dt + T(PI_PON, n) * kPON * dt;
            *F[PI_PON] = *F[PI_PON] - T(PI_PON, n) * kPON * dt;
            *F[PI_PO4] = *F[PI_PO4] + T(PI_DOP, n) * kDOP * dt;
            *F[PI_DOP] = *F[PI_DOP] - T(PI_DOP, n) * kDOP * dt + T(PI_POP, n) * kPOP * dt;
            *F[PI_POP] = *F[PI_POP] - T(PI_POP, n) * kPOP * dt;
            FT DFe = T(PI_DFe, n);
            FT excess = M::max((DFe - lig), z0);
            *F[PI_DFe] = *F[PI_DFe] + T(PI_PFe_bio, n) * kPOC * dt -
                         (lPOC * T(PI_POC, n) + lmin + ldust * T(PI_Dust, n)) * DFe * DFeFrac * dt +
                         T(PI_PFe_inorg, n) * kdiss * dt -
                         excess * DFe * DFeFrac * lFe * dt;
            *F[PI_PFe_bio] = *F[PI_PFe_bio] + T(PI_POC, n) * DFe * DFeFrac * lPOC * dt -
                             kPOC * T(PI_PFe_bio, n) * dt;
            *F[PI_PFe_inorg] = *F[PI_PFe_inorg] - T(PI_PFe_inorg, n) * kdiss * dt +
                               (lmin + ldust * T(PI_Dust, n)) * DFe * DFeFrac * dt +
                               excess * DFe * DFeFrac * lFe * dt;
        }
    }
    /* getbc (HaloRegions/boundary_conditions.jl:31-40): bc, bc[a, b] or bc[a, b, t] */
    FT getbc(int q, int f, int a, int b, int t) const {
        if (bck[q][f] == PI_BCK_CONST) return bcv[q][f];
        const int64_t na = f <= PI_BC_EAST ? Ny : Nx, nb = f <= PI_BC_NORTH ? Nz : Ny;
        const int64_t tt = bck[q][f] == PI_BCK_SERIES ? (int64_t)(t - 1) : 0;
        return bca[q][f][(int64_t)(a - 1) + na * ((int64_t)(b - 1) + nb * tt)];
    }
    /* apply_bcs! (HaloRegions/apply_bcs.jl:5-99): per tracer the west, east, south,
     * north, bottom and top kernels in that order, each a no-op without a condition */
    void apply_bcs(int iter, FT dt) {
        for (int q = 0; q < PI_NTRACERS; ++q) {
            V& G = Gcs[q];
            if (bck[q][PI_BC_WEST])
                for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) {
                    const int ii = 1 + Hx, jj = j + Hy, kk = k + Hz;
                    G[I(ii, jj, kk)] += getbc(q, PI_BC_WEST, j, k, iter) * dt * Ax(ii, jj, kk) / vol(ii, jj, kk);
                }
            if (bck[q][PI_BC_EAST])
                for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j) {
                    const int ii = Nx + Hx, jj = j + Hy, kk = k + Hz;
                    G[I(ii, jj, kk)] -= getbc(q, PI_BC_EAST, j, k, iter) * dt * Ax(ii, jj, kk) / vol(ii, jj, kk);
                }
            if (bck[q][PI_BC_SOUTH])
                for (int k = 1; k <= Nz; ++k) for (int i = 1; i <= Nx; ++i) {
                    const int ii = i + Hx, jj = 1 + Hy, kk = k + Hz;
                    G[I(ii, jj, kk)] += getbc(q, PI_BC_SOUTH, i, k, iter) * dt * Ay(ii, jj, kk) / vol(ii, jj, kk);
                }
            if (bck[q][PI_BC_NORTH])
                for (int k = 1; k <= Nz; ++k) for (int i = 1; i <= Nx; ++i) {
                    const int ii = i + Hx, jj = Ny + Hy, kk = k + Hz;
                    G[I(ii, jj, kk)] -= getbc(q, PI_BC_NORTH, i, k, iter) * dt * Ay(ii, jj, kk) / vol(ii, jj, kk);
                }
            if (bck[q][PI_BC_BOTTOM])
                for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
                    const int ii = i + Hx, jj = j + Hy, kk = Nz + Hz;
                    G[I(ii, jj, kk)] += getbc(q, PI_BC_BOTTOM, i, j, iter) * dt * Az(ii, jj) / vol(ii, jj, kk);
                }
            if (bck[q][PI_BC_TOP])
                for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i) {
                    const int ii = i + Hx, jj = j + Hy, kk = 1 + Hz;
                    G[I(ii, jj, kk)] -= getbc(q, PI_BC_TOP, i, j, iter) * dt * Az(ii, jj) / vol(ii, jj, kk);
                }
        }
    }
    /* tracer_update! (Biogeochemistry/tracer_update.jl:22-43) */
    void tracer_update(FT dt) {
        tracer_advection(dt);
        tracer_diffusion(dt);
        tracer_sinking(dt);
        for (int q = 0; q < PI_NTRACERS; ++q) std::fill(ttemp[q].begin(), ttemp[q].end(), (FT)0);
        tracer_forcing(dt);
        apply_bcs((int)iteration, dt);
        for (int q = 0; q < PI_NTRACERS; ++q)       /* apply_tendency_kernel! (:6-20) */
            OMP_FOR
            for (int k = Hz + 1; k <= Hz + Nz; ++k) for (int j = Hy + 1; j <= Hy + Ny; ++j)
                for (int i = Hx + 1; i <= Hx + Nx; ++i) {
                    int64_t n = I(i, j, k);
                    tr[q][n] = tr[q][n] + (Gcs[q][n] + plk[q][n] / vol(i, j, k));
                }
        fill_halo_tracer(tr);
    }

    /* particles_from_bcs! (Individuals/Abiotic/particle_interaction.jl:158-267):
     * tr_temp = 0; calc_particle_bc_{top,bottom,east,west,north,south}! write
     * isless(rnd·sz_min·Nsuper, bc·ΔT·A) at the cells next to each face with a
     * condition (later faces overwrite earlier ones); the k-th flagged cell in
     * column-major order fills the k-th free slot at (i, j, k) - H - 0.4. */
    pi_status particles_from_bcs(int a, FT dt) {
        Plank& p = ab[a];
        const FT szN = AP[a][PI_A_sz_min], Ns = AP[a][PI_A_Nsuper];
        V flag(ncell, (FT)0);
        auto rnd = [&](int64_t c) {
            float u0, u1, u2;
            pi_uniform3(seed, iteration, PI_ABIO_SPECIES + a, PI_KIND_ABIO_BC, c, &u0, &u1, &u2);
            return (FT)u0;
        };
        auto bcv = [&](int f, int i, int j) -> FT {
            if (abck[a][f] == PI_BCK_CONST) return abcv[a][f];
            const int64_t na = f <= PI_BC_EAST ? Ny : Nx, nb = f <= PI_BC_NORTH ? Nz : Ny;
            const int64_t tt = abck[a][f] == PI_BCK_SERIES ? iteration - 1 : 0;
            return abca[a][f][(int64_t)(i - 1) + na * ((int64_t)(j - 1) + nb * tt)];
        };
        auto put = [&](int ii, int jj, int kk, FT b, FT A) {
            const int64_t c = I(ii, jj, kk);
            flag[c] = M::isless(rnd(c) * szN * Ns, b * dt * A) ? (FT)1 : (FT)0;
        };
        if (abck[a][PI_BC_TOP])
            for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i)
                put(i + Hx, j + Hy, 1 + Hz, bcv(PI_BC_TOP, i, j), Az(i + Hx, j + Hy));
        if (abck[a][PI_BC_BOTTOM])
            for (int j = 1; j <= Ny; ++j) for (int i = 1; i <= Nx; ++i)
                put(i + Hx, j + Hy, Nz + Hz, bcv(PI_BC_BOTTOM, i, j), Az(i + Hx, j + Hy));
        if (abck[a][PI_BC_EAST])
            for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j)
                put(Nx + Hx, j + Hy, k + Hz, bcv(PI_BC_EAST, j, k), Ax(Nx + Hx, j + Hy, k + Hz));
        if (abck[a][PI_BC_WEST])
            for (int k = 1; k <= Nz; ++k) for (int j = 1; j <= Ny; ++j)
                put(1 + Hx, j + Hy, k + Hz, bcv(PI_BC_WEST, j, k), Ax(1 + Hx, j + Hy, k + Hz));
        if (abck[a][PI_BC_NORTH])
            for (int k = 1; k <= Nz; ++k) for (int i = 1; i <= Nx; ++i)
                put(i + Hx, Ny + Hy, k + Hz, bcv(PI_BC_NORTH, i, k), Ay(i + Hx, Ny + Hy, k + Hz));
        if (abck[a][PI_BC_SOUTH])
            for (int k = 1; k <= Nz; ++k) for (int i = 1; i <= Nx; ++i)
                put(i + Hx, 1 + Hy, k + Hz, bcv(PI_BC_SOUTH, i, k), Ay(i + Hx, 1 + Hy, k + Hz));
        std::vector<int64_t> inds, freei;
        for (int64_t c = 0; c < ncell; ++c) if (flag[c] == (FT)1) inds.push_back(c);    /* findall, column-major */
        if (inds.empty()) return PI_OK;
        for (int64_t i = 0; i < cap; ++i) if (!p.ac[i]) freei.push_back(i);
        if (inds.size() > freei.size()) {
            failed = true;
            return fail(PI_ERR_CAPACITY, "number of abiotic particles exceeds the capacity");
        }
        for (size_t k = 0; k < inds.size(); ++k) {       /* copy_abiotic_particle_from_field_kernel! */
            const int64_t c = inds[k], r = freei[k];
            const int i = (int)(c % Px) + 1, j = (int)((c / Px) % Py) + 1, kk = (int)(c / ((int64_t)Px * Py)) + 1;
            p.ac[r] = 1;
            p.x[r] = (FT)(float)(i - Hx) - L(0.4f);
            p.y[r] = (FT)(float)(j - Hy) - L(0.4f);
            p.z[r] = (FT)(float)(kk - Hz) - L(0.4f);
        }
        return PI_OK;
    }

    /* ================================================== TimeStep! =======
     * Model/time_step.jl:12-180 */
    pi_status step(float dt_in) {
        if (failed) return fail(PI_ERR_CAPACITY, "model is in an error state from an earlier step");
        for (int q = 0; q < PI_NTRACERS; ++q)      /* bc[:, :, iteration] must exist */
            for (int f = 0; f < PI_NBC; ++f)
                if (bcnt[q][f] && iteration + 1 > bcnt[q][f])
                    return fail(PI_ERR_ARG, "BC: time-dependent flux boundary condition shorter than the run");
        for (int a = 0; a < nsa; ++a)
            for (int f = 0; f < PI_NBC; ++f)
                if (abcnt[a][f] && iteration + 1 > abcnt[a][f])
                    return fail(PI_ERR_ARG, "BC: time-dependent particle boundary condition shorter than the run");
        const FT dt = (FT)dt_in;
        for (const auto& p : sp) for (int64_t i = 0; i < cap; ++i) ind_steps += p.ac[i] ? 1 : 0;
        iteration += 1;
        t = (FT)(float)iteration * dt;
        for (int q = 0; q < 3; ++q)
            OMP_FOR
            for (int64_t n = 0; n < ncell; ++n) velh[q][n] = (vel0[q][n] + vel1[q][n]) * L(0.5f);
        for (int q = 0; q < PI_NTRACERS; ++q) std::fill(plk[q].begin(), plk[q].end(), (FT)0);
        std::fill(Chl.begin(), Chl.end(), (FT)0);
        for (int a = 0; a < nsa; ++a) {                 /* abiotic particles (time_step.jl:21-38) */
            pi_status st = particles_from_bcs(a, dt);
            if (st != PI_OK) return st;
            particle_advection(ab[a], dt);
            particle_diffusion(ab[a], PI_ABIO_SPECIES + a, dt);
            find_inds(ab[a]);
        }
        if (B[PI_B_shared_graz] == L(1)) {
            std::fill(pop.begin(), pop.end(), (FT)0);
            for (int s = 0; s < nsp; ++s) {
                particle_advection(sp[s], dt);
                particle_diffusion(sp[s], s, dt);
                find_inds(sp[s]);
                acc_chl(sp[s]);
                acc_counts(sp[s]);
            }
            for (int ki = 1; ki <= Nz; ++ki) calc_par(ki);
            for (int s = 0; s < nsp; ++s) {
                find_NPT(sp[s]);
                pi_status st = plankton_update(s, dt);
                if (st != PI_OK) return st;
            }
        } else {
            for (int s = 0; s < nsp; ++s) {
                particle_advection(sp[s], dt);
                particle_diffusion(sp[s], s, dt);
                find_inds(sp[s]);
                acc_chl(sp[s]);
            }
            for (int ki = 1; ki <= Nz; ++ki) calc_par(ki);
            for (int s = 0; s < nsp; ++s) {
                std::fill(pop.begin(), pop.end(), (FT)0);
                acc_counts(sp[s]);
                find_NPT(sp[s]);
                pi_status st = plankton_update(s, dt);
                if (st != PI_OK) return st;
            }
        }
        for (const auto& pr : pint) particle_interaction(pr[0], pr[1]);   /* time_step.jl:120-128 */
        for (size_t q = 0; q < prel.size(); ++q) {                        /* :130-137 */
            pi_status st = particle_release(prel[q][0], prel[q][1], (int)q, dt);
            if (st != PI_OK) return st;
        }
        if (nsa > 0 && mode == PI_MODE_IRONENERGY)       /* diags_proc!(ptc, ptc, ac, ...) (:140-149) */
            for (int s = 0; s < nsp; ++s)
                for (int64_t i = 0; i < cap; ++i) diagSp[s][4][cell_of(sp[s], i)] += M::mask(sp[s].ac[i], sp[s].ptc[i]);
        for (int a = 0; a < nsa; ++a)                   /* diags_proc!(num, ac, ac, xi, yi, zi) (:152-159) */
            for (int64_t i = 0; i < cap; ++i) abNum[a][cell_of(ab[a], i)] += (FT)(ab[a].ac[i] ? 1 : 0);
        tracer_update(dt);
        OMP_FOR
        for (int64_t n = 0; n < ncell; ++n) diagPAR[n] += par[n];
        for (size_t a = 0; a < diag_tr.size(); ++a)
            OMP_FOR
            for (int64_t n = 0; n < ncell; ++n) diagTr[a][n] += tr[diag_tr[a]][n];
        for (int q = 0; q < 3; ++q) vel0[q] = vel1[q];
        par0 = par;
        return PI_OK;
    }
};

/* ====================================================== C ABI wrappers == */
template <typename FT> static Oracle<FT>* as(orc_model* m) { return static_cast<Oracle<FT>*>(m); }

#define DISPATCH(m, expr)                                   \
    do {                                                    \
        if (!(m)) return fail(PI_ERR_ARG, "null model");    \
        if ((m)->precision == 64) { auto* o = as<double>(m); expr; } \
        else { auto* o = as<float>(m); expr; }              \
    } while (0)

const char* orc_last_error(void) { return g_err.c_str(); }

pi_status orc_model_create(const pi_model_desc* d, int32_t precision, orc_model** out) {
    if (!d || !out) return fail(PI_ERR_ARG, "null argument");
    if (precision == 64) {
        auto* o = new Oracle<double>();
        o->precision = 64;
        pi_status s = o->init(d);
        if (s != PI_OK) { delete o; return s; }
        *out = o;
    } else {
        auto* o = new Oracle<float>();
        o->precision = 32;
        pi_status s = o->init(d);
        if (s != PI_OK) { delete o; return s; }
        *out = o;
    }
    return PI_OK;
}

pi_status orc_model_destroy(orc_model* m) { delete m; return PI_OK; }

template <typename FT>
static pi_status upload_impl(Oracle<FT>* o, int32_t s, const pi_individuals_host* h) {
    if (s < 0 || s >= o->nsp) return fail(PI_ERR_ARG, "species out of range");
    auto& p = o->sp[s];
    auto cp = [&](std::vector<FT>& dst, const float* src) {
        for (int64_t i = 0; i < o->cap; ++i) dst[i] = src ? (FT)src[i] : (FT)0;
    };
    cp(p.x, h->x); cp(p.y, h->y); cp(p.z, h->z); cp(p.Sz, h->Sz); cp(p.Bm, h->Bm); cp(p.Cq, h->Cq);
    cp(p.Nq, h->Nq); cp(p.Pq, h->Pq); cp(p.Chl, h->Chl); cp(p.gen, h->gen); cp(p.age, h->age);
    cp(p.graz, h->graz); cp(p.mort, h->mort); cp(p.dvid, h->dvid);
    if (o->mode == PI_MODE_MACROMOLECULAR) cp(p.RNA, h->ex[0]);
    if (o->mode == PI_MODE_IRONENERGY) {
        std::vector<FT>* dst[8] = {&p.qNO3, &p.qFe, &p.qFePS, &p.qFeNR, &p.qFeNF, &p.tdark, &p.exEn, &p.ptc};
        for (int k = 0; k < 8; ++k) cp(*dst[k], h->ex[k]);
    }
    for (int64_t i = 0; i < o->cap; ++i) {
        p.xi[i] = h->xi ? h->xi[i] : 0; p.yi[i] = h->yi ? h->yi[i] : 0; p.zi[i] = h->zi ? h->zi[i] : 0;
        p.ac[i] = h->ac ? (h->ac[i] != 0) : 0;
    }
    return PI_OK;
}
pi_status orc_upload_individuals(orc_model* m, int32_t sp, const pi_individuals_host* h) {
    if (!h) return fail(PI_ERR_ARG, "null host view");
    DISPATCH(m, return upload_impl(o, sp, h));
}
pi_status orc_upload_tracer(orc_model* m, int32_t q, const float* c) {
    if (q < 0 || q >= PI_NTRACERS || !c) return fail(PI_ERR_ARG, "tracer id");
    DISPATCH(m, for (int64_t n = 0; n < o->ncell; ++n) o->tr[q][n] = c[n]; return PI_OK);
}
pi_status orc_set_velocity(orc_model* m, const float* u, const float* v, const float* w) {
    if (!u || !v || !w) return fail(PI_ERR_ARG, "null velocity");
    DISPATCH(m, o->set_velocity(u, v, w); return PI_OK);
}
pi_status orc_set_parf(orc_model* m, const float* parf) {
    if (!parf) return fail(PI_ERR_ARG, "null PARF");
    DISPATCH(m, for (int64_t n = 0; n < (int64_t)o->Nx * o->Ny; ++n) o->PARF[n] = parf[n]; return PI_OK);
}
pi_status orc_set_temperature(orc_model* m, const float* tt) {
    if (!tt) return fail(PI_ERR_ARG, "null temperature");
    DISPATCH(m, o->set_temperature(tt); return PI_OK);
}
pi_status orc_step(orc_model* m, float dt) { DISPATCH(m, return o->step(dt)); }
template <typename FT>
static pi_status set_bc_impl(Oracle<FT>* o, int32_t q, int32_t f, int32_t kind, const float* v, int64_t nt) {
    if (q < 0 || q >= PI_NTRACERS) return fail(PI_ERR_ARG, "tracer id");
    if (f < 0 || f >= PI_NBC) return fail(PI_ERR_ARG, "boundary position");
    if (kind < PI_BCK_NONE || kind > PI_BCK_SERIES) return fail(PI_ERR_ARG, "boundary condition kind");
    if (kind != PI_BCK_NONE && !v) return fail(PI_ERR_ARG, "null boundary condition values");
    if (kind == PI_BCK_SERIES && nt < 1) return fail(PI_ERR_ARG, "BC: nt >= 1 required");
    const int64_t n1 = f <= PI_BC_EAST ? o->Ny : o->Nx, n2 = f <= PI_BC_NORTH ? o->Nz : o->Ny;
    const int64_t n = kind == PI_BCK_FIELD ? n1 * n2 : kind == PI_BCK_SERIES ? n1 * n2 * nt : 0;
    o->bck[q][f] = kind;
    o->bcv[q][f] = kind == PI_BCK_CONST ? (FT)v[0] : (FT)0;     /* FT.(bc_value) */
    o->bca[q][f].assign(n, (FT)0);
    for (int64_t a = 0; a < n; ++a) o->bca[q][f][a] = (FT)v[a];
    o->bcnt[q][f] = kind == PI_BCK_SERIES ? nt : 0;
    return PI_OK;
}
pi_status orc_set_tracer_bc(orc_model* m, int32_t q, int32_t f, int32_t kind, const float* v, int64_t nt) {
    DISPATCH(m, return set_bc_impl(o, q, f, kind, v, nt));
}
template <typename FT>
static pi_status set_pbc_impl(Oracle<FT>* o, int32_t a, int32_t f, int32_t kind, const float* v, int64_t nt) {
    if (a < 0 || a >= o->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
    if (f < 0 || f >= PI_NBC) return fail(PI_ERR_ARG, "boundary position");
    if (kind < PI_BCK_NONE || kind > PI_BCK_SERIES) return fail(PI_ERR_ARG, "boundary condition kind");
    if (kind != PI_BCK_NONE && !v) return fail(PI_ERR_ARG, "null boundary condition values");
    if (kind == PI_BCK_SERIES && nt < 1) return fail(PI_ERR_ARG, "BC: nt >= 1 required");
    const int64_t n1 = f <= PI_BC_EAST ? o->Ny : o->Nx, n2 = f <= PI_BC_NORTH ? o->Nz : o->Ny;
    const int64_t n = kind == PI_BCK_FIELD ? n1 * n2 : kind == PI_BCK_SERIES ? n1 * n2 * nt : 0;
    o->abck[a][f] = kind;
    o->abcv[a][f] = kind == PI_BCK_CONST ? (FT)v[0] : (FT)0;
    o->abca[a][f].assign(n, (FT)0);
    for (int64_t i = 0; i < n; ++i) o->abca[a][f][i] = (FT)v[i];
    o->abcnt[a][f] = kind == PI_BCK_SERIES ? nt : 0;
    return PI_OK;
}
pi_status orc_set_particle_bc(orc_model* m, int32_t a, int32_t f, int32_t kind, const float* v, int64_t nt) {
    DISPATCH(m, return set_pbc_impl(o, a, f, kind, v, nt));
}
template <typename FT>
static pi_status up_abio_impl(Oracle<FT>* o, int32_t a, const float* x, const float* y, const float* z, const uint8_t* ac) {
    if (a < 0 || a >= o->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
    auto& p = o->ab[a];
    for (int64_t i = 0; i < o->cap; ++i) {
        p.x[i] = (FT)x[i]; p.y[i] = (FT)y[i]; p.z[i] = (FT)z[i]; p.ac[i] = ac[i] ? 1 : 0;
        p.xi[i] = p.yi[i] = p.zi[i] = 0;
    }
    return PI_OK;
}
pi_status orc_upload_abiotic(orc_model* m, int32_t a, const float* x, const float* y, const float* z, const uint8_t* ac) {
    if (!x || !y || !z || !ac) return fail(PI_ERR_ARG, "null argument");
    DISPATCH(m, return up_abio_impl(o, a, x, y, z, ac));
}
template <typename FT>
static pi_status down_abio_impl(Oracle<FT>* o, int32_t a, int32_t f, void* h, int64_t n) {
    if (a < 0 || a >= o->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
    if (n != o->cap) return fail(PI_ERR_ARG, "n must equal capacity");
    const auto& p = o->ab[a];
    for (int64_t i = 0; i < n; ++i) {
        switch (f) {
            case PI_AB_x: ((float*)h)[i] = (float)p.x[i]; break;
            case PI_AB_y: ((float*)h)[i] = (float)p.y[i]; break;
            case PI_AB_z: ((float*)h)[i] = (float)p.z[i]; break;
            case PI_AB_xi: ((int64_t*)h)[i] = p.xi[i]; break;
            case PI_AB_yi: ((int64_t*)h)[i] = p.yi[i]; break;
            case PI_AB_zi: ((int64_t*)h)[i] = p.zi[i]; break;
            case PI_AB_ac: ((uint8_t*)h)[i] = p.ac[i]; break;
            default: return fail(PI_ERR_ARG, "abiotic field id");
        }
    }
    return PI_OK;
}
pi_status orc_download_abiotic(orc_model* m, int32_t a, int32_t f, void* h, int64_t n) {
    if (!h) return fail(PI_ERR_ARG, "null argument");
    DISPATCH(m, return down_abio_impl(o, a, f, h, n));
}
pi_status orc_download_abiotic_diag(orc_model* m, int32_t a, float* h) {
    if (!h) return fail(PI_ERR_ARG, "null argument");
    DISPATCH(m, if (a < 0 || a >= o->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
             for (int64_t n = 0; n < o->ncell; ++n) h[n] = (float)o->abNum[a][n]; return PI_OK);
}

template <typename FT, typename OUT>
static pi_status field_impl(Oracle<FT>* o, int32_t s, int32_t f, OUT* out, int64_t n, bool ints_ok) {
    if (s < 0 || s >= o->nsp) return fail(PI_ERR_ARG, "species out of range");
    if (n != o->cap) return fail(PI_ERR_ARG, "n must equal capacity");
    auto& p = o->sp[s];
    const std::vector<FT>* src = nullptr;
    switch (f) {
        case PI_F_x: src = &p.x; break; case PI_F_y: src = &p.y; break; case PI_F_z: src = &p.z; break;
        case PI_F_Sz: src = &p.Sz; break; case PI_F_Bm: src = &p.Bm; break; case PI_F_Cq: src = &p.Cq; break;
        case PI_F_Nq: src = &p.Nq; break; case PI_F_Pq: src = &p.Pq; break; case PI_F_Chl: src = &p.Chl; break;
        case PI_F_gen: src = &p.gen; break; case PI_F_age: src = &p.age; break;
        case PI_F_graz: src = &p.graz; break; case PI_F_mort: src = &p.mort; break; case PI_F_dvid: src = &p.dvid; break;
        case PI_F_ex0: if (o->mode == PI_MODE_MACROMOLECULAR) src = &p.RNA; break;   /* PI_F_RNA */
        default: break;
    }
    if (o->mode == PI_MODE_IRONENERGY && f >= PI_F_ex0 && f < PI_F_ex0 + 8) {
        const std::vector<FT>* ie[8] = {&p.qNO3, &p.qFe, &p.qFePS, &p.qFeNR, &p.qFeNF, &p.tdark, &p.exEn, &p.ptc};
        src = ie[f - PI_F_ex0];
    }
    if (src) { for (int64_t i = 0; i < n; ++i) out[i] = (OUT)(*src)[i]; return PI_OK; }
    if (!ints_ok) return fail(PI_ERR_ARG, "field is not floating point");
    if (f == PI_F_xi || f == PI_F_yi || f == PI_F_zi) {
        const auto& v = f == PI_F_xi ? p.xi : f == PI_F_yi ? p.yi : p.zi;
        int64_t* o64 = reinterpret_cast<int64_t*>(out);
        for (int64_t i = 0; i < n; ++i) o64[i] = v[i];
        return PI_OK;
    }
    if (f == PI_F_ac) {
        uint8_t* o8 = reinterpret_cast<uint8_t*>(out);
        for (int64_t i = 0; i < n; ++i) o8[i] = p.ac[i];
        return PI_OK;
    }
    return fail(PI_ERR_ARG, "unknown field");
}
pi_status orc_download_field(orc_model* m, int32_t sp, int32_t f, void* host, int64_t n) {
    DISPATCH(m, return field_impl(o, sp, f, reinterpret_cast<float*>(host), n, true));
}
pi_status orc_download_field_f64(orc_model* m, int32_t sp, int32_t f, double* host, int64_t n) {
    DISPATCH(m, return field_impl(o, sp, f, host, n, false));
}
pi_status orc_download_tracer(orc_model* m, int32_t q, float* c) {
    if (q < 0 || q >= PI_NTRACERS) return fail(PI_ERR_ARG, "tracer id");
    DISPATCH(m, for (int64_t n = 0; n < o->ncell; ++n) c[n] = (float)o->tr[q][n]; return PI_OK);
}
pi_status orc_download_tracer_f64(orc_model* m, int32_t q, double* c) {
    if (q < 0 || q >= PI_NTRACERS) return fail(PI_ERR_ARG, "tracer id");
    DISPATCH(m, for (int64_t n = 0; n < o->ncell; ++n) c[n] = (double)o->tr[q][n]; return PI_OK);
}
template <typename FT>
static pi_status grid_impl(Oracle<FT>* o, int32_t which, float* h, int64_t n) {
    const std::vector<FT>* src = nullptr;
    switch (which) {
        case PI_G_par: src = &o->par; break; case PI_G_par0: src = &o->par0; break;
        case PI_G_Chl: src = &o->Chl; break; case PI_G_pop: src = &o->pop; break;
        case PI_G_temp: src = &o->temp; break; case PI_G_PARF: src = &o->PARF; break;
        case PI_G_u: src = &o->vel1[0]; break; case PI_G_v: src = &o->vel1[1]; break; case PI_G_w: src = &o->vel1[2]; break;
        case PI_G_diag_PAR: src = &o->diagPAR; break;
        default:
            if (which >= PI_G_plk_DIC && which < PI_G_plk_DIC + PI_NTRACERS) src = &o->plk[which - PI_G_plk_DIC];
    }
    if (!src) return fail(PI_ERR_ARG, "unknown grid field");
    if (n != (int64_t)src->size()) return fail(PI_ERR_ARG, "size mismatch");
    for (int64_t i = 0; i < n; ++i) h[i] = (float)(*src)[i];
    return PI_OK;
}
pi_status orc_download_grid(orc_model* m, int32_t which, float* h, int64_t n) {
    DISPATCH(m, return grid_impl(o, which, h, n));
}
pi_status orc_download_diag(orc_model* m, int32_t s, int32_t which, float* h) {
    DISPATCH(m, if (s < 0 || s >= o->nsp || which < 0 || which >= PI_NDIAG) return fail(PI_ERR_ARG, "diag id");
             for (int64_t n = 0; n < o->ncell; ++n) h[n] = (float)o->diagSp[s][which][n]; return PI_OK);
}
pi_status orc_download_diag_tracer(orc_model* m, int32_t q, float* h) {
    DISPATCH(m, for (size_t a = 0; a < o->diag_tr.size(); ++a) if (o->diag_tr[a] == q) {
                    for (int64_t n = 0; n < o->ncell; ++n) h[n] = (float)o->diagTr[a][n];
                    return PI_OK; }
             return fail(PI_ERR_ARG, "tracer not diagnosed"));
}
pi_status orc_reset_diags(orc_model* m) {
    DISPATCH(m, std::fill(o->diagPAR.begin(), o->diagPAR.end(), 0);
             for (auto& v : o->diagTr) std::fill(v.begin(), v.end(), 0);
             for (auto& a : o->diagSp) for (auto& v : a) std::fill(v.begin(), v.end(), 0);
             for (auto& v : o->abNum) std::fill(v.begin(), v.end(), 0);
             return PI_OK);
}
pi_status orc_set_diag_tracers(orc_model* m, int32_t n, const int32_t* ids) {
    for (int32_t a = 0; a < n; ++a)
        if (ids[a] < 0 || ids[a] >= PI_NTRACERS) return fail(PI_ERR_ARG, "tracer id");
    DISPATCH(m, o->diag_tr.assign(ids, ids + n);
             o->diagTr.assign(o->diag_tr.size(), std::vector<decltype(o->t)>(o->ncell, 0)); return PI_OK);
}
pi_status orc_get_iteration(orc_model* m, int64_t* it) { DISPATCH(m, *it = o->iteration; return PI_OK); }
pi_status orc_individual_steps(orc_model* m, int64_t* n) { DISPATCH(m, *n = o->ind_steps; return PI_OK); }
pi_status orc_count_active(orc_model* m, int32_t s, int64_t* n) {
    DISPATCH(m, if (s < 0 || s >= o->nsp) return fail(PI_ERR_ARG, "species");
             int64_t c = 0; for (auto a : o->sp[s].ac) c += a; *n = c; return PI_OK);
}

pi_status orc_fill_halo(int32_t Nx, int32_t Ny, int32_t Nz, int32_t Hx, int32_t Hy, int32_t Hz,
                        int32_t tx, int32_t ty, int32_t tz, float* c) {
    Oracle<float> o;
    o.Nx = Nx; o.Ny = Ny; o.Nz = Nz; o.Hx = Hx; o.Hy = Hy; o.Hz = Hz;
    o.Px = Nx + 2 * Hx; o.Py = Ny + 2 * Hy; o.Pz = Nz + 2 * Hz;
    o.tx = tx; o.ty = ty; o.tz = tz;
    o.fill_halo_generic(c, tx, ty);
    return PI_OK;
}
pi_status orc_fill_halo_op(int32_t op, int32_t dim, int32_t topo, int32_t Nx, int32_t Ny, int32_t Nz,
                           int32_t Hx, int32_t Hy, int32_t Hz, float* c)
{
    if (dim < 0 || dim > 2 || op < 0 || op > 5) return fail(PI_ERR_ARG, "halo op");
    Oracle<float> o;
    o.Nx = Nx; o.Ny = Ny; o.Nz = Nz; o.Hx = Hx; o.Hy = Hy; o.Hz = Hz;
    o.Px = Nx + 2 * Hx; o.Py = Ny + 2 * Hy; o.Pz = Nz + 2 * Hz;
    o.tx = o.ty = o.tz = topo;
    switch (op) {
        case 0: o.fill_lo(c, dim); break;
        case 1: o.fill_hi(c, dim); break;
        case 2: o.fill_hi_vel(c, dim); break;
        case 3: o.fill_hi_Gc(c, dim); break;
        case 4: o.fill_top_flux_sink(c); break;
        default: o.fill_bottom_flux_sink(c); break;
    }
    return PI_OK;
}
int32_t orc_set_threads(int32_t n)
{
    g_threads = n < 1 ? 1 : n;
    return g_threads;
}
float orc_expf(float x) { return pi_expf(x); }
double orc_exp_f64(double x) { return pi_exp_f64(x); }
void orc_log_f64_v(const double* x, double* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = pi_log_f64(x[i]); }
void orc_cos2pi_f64_v(const double* x, double* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = pi_cos2pi_f64(x[i]); }
void orc_init_normal_v(const float* a, const float* b, float* y, int64_t n) { for (int64_t i = 0; i < n; ++i) y[i] = pi_init_normal(a[i], b[i]); }
void orc_uniform3(uint64_t seed, int64_t it, int32_t s, int32_t kind, int64_t slot, float* o3) {
    pi_uniform3(seed, it, s, kind, slot, &o3[0], &o3[1], &o3[2]);
}
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { pi_philox4x32_10(ctr, key, out); }
float orc_tempfunc(const float* params, float T) {
    std::array<float, PI_NPARAMS> q; for (int i = 0; i < PI_NPARAMS; ++i) q[i] = params[i];
    return Oracle<float>::tempFunc(T, q);
}
float orc_tempfunc_ps(const float* params, float T) {
    std::array<float, PI_NPARAMS> q; for (int i = 0; i < PI_NPARAMS; ++i) q[i] = params[i];
    return Oracle<float>::tempFunc_PS(T, q);
}

