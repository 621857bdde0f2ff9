/*
 * pi_host.hip - the C ABI of libpi_hip.so (include/pi_hip.h): model storage in
 * HBM, input staging, the per-step launch sequence and the reference-layout
 * downloads.
 *
 * One step (Model/time_step.jl:12-180) is this fixed launch sequence on the
 * model's stream, with no host synchronisation:
 *   [vel½]  marks -> bitmap scan -> capacity check
 *   adv -> cell-count scan -> place -> cells (+bigcells) -> par -> phys -> plk
 *   sweep x -> fill y -> sweep y -> fill z -> final -> fill x,y,z -> [diag tracers]
 *   finish (flip current buffers, count the store, iteration += 1)
 */
#include "pi_device.h"

#include <cstdio>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>
#include <cstddef>
#include <mutex>

#include "k_args.h"

static thread_local std::string g_err;
static pi_status fail(pi_status s, const std::string& m) { g_err = m; return s; }
static pi_status scratch_base(int dev, char** out);
#define HIPCHECK(x)                                                                      \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(PI_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

/* One step of one model as a plan: the fixed sequence of kernel launches, each
 * with its argument block and block count (0 = skipped for this model).  Every
 * model's plan has the same sequence of kernels, so the plans of many models
 * launch position by position as batched kernels (pi_batch_*). */
struct Rec { uint16_t kid; uint8_t side; uint32_t nblk; uint32_t off; const char* grp; };
/* The launches of a step.  Records between a K_FORK and a K_JOIN marker that
 * were added in side mode run on a second stream (they depend only on what
 * was ready at the fork), the others stay on the model's stream; the join
 * makes the model's stream wait for the side stream. */
struct Plan {
    std::vector<Rec> recs;
    std::vector<uint8_t> args;
    bool side_mode = false;
    void mark(int k) { recs.push_back(Rec{(uint16_t)k, 0, 0, (uint32_t)args.size(), "mark"}); }
    void side(bool on) { side_mode = on; }
    template <class A> void add(KId k, const char* grp, long long nblk, const A& a)
    {
        Rec r;
        r.kid = (uint16_t)k;
        r.side = side_mode ? 1 : 0;
        r.nblk = (uint32_t)std::max(0ll, nblk);
        r.off = (uint32_t)args.size();
        r.grp = grp;
        args.resize(args.size() + ((sizeof(A) + 15) & ~size_t(15)), 0);
        std::memcpy(args.data() + r.off, &a, sizeof(A));
        recs.push_back(r);
    }
    void clear() { recs.clear(); args.clear(); side_mode = false; }
};
template <class A> static A zeroed() { A a; std::memset((void*)&a, 0, sizeof(A)); return a; }

static KInfo g_kinfo[K_NKID];
static const KInfo* kinfo()
{
    static bool done = false;
    if (!done) { reg_particles(g_kinfo); reg_tracers(g_kinfo); reg_util(g_kinfo); done = true; }
    return g_kinfo;
}

struct KTimer { const char* name; hipEvent_t a, b; };

struct pi_model {
    int dev = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;          /* second stream of the step's fork (default priority) */
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    DevGrid g{};
    int nsp = 0;
    long long cap = 0, ctot = 0, ncell = 0, nint = 0;
    int nw = 0;
    std::vector<float> h_params, bgc;
    float* d_params = nullptr;
    uint64_t seed = 0;
    DevState* d_state = nullptr;
    DevState h_state{};
    SlotTable tab{};
    float* T1[PI_NTRACERS]{};
    float* T2[PI_NTRACERS]{};
    float* vel[3][3]{};        /* buffer set b, component c */
    int v0 = 0, v1 = 0;        /* current buffer sets of vel₀ and vel₁ */
    bool vel_same = true;
    float *Chl = nullptr, *pop = nullptr, *temp = nullptr, *PARF_in = nullptr, *PARF_out = nullptr;
    float* plk[PI_NTRACERS]{};  /* null for the tracers without particle exchange */
    float* plk_slab = nullptr;
    float* diagPAR = nullptr;
    float* diag = nullptr;     /* [sp][4][ncell] */
    std::vector<int> diag_ids;
    std::vector<float*> diag_tr;
    uint32_t *cnt = nullptr, *seg = nullptr, *partials = nullptr, *pkey = nullptr, *psrc = nullptr;
    uint32_t *skey = nullptr, *ssrc = nullptr, *big = nullptr, *lng = nullptr, *bm = nullptr, *G = nullptr;
    int long_blocks = 1;       /* blocks of the one-block-per-long-cell kernels */
    int ncu = 256;             /* compute units of the device */
    uint32_t qcap = 0;         /* entries per queue shard */
    Plan plan;                 /* the launches of the last step built */
    bool dry = false;          /* create_impl only sums the allocation sizes */
    long long bytes = 0;       /* device bytes allocated (256-byte granules) */
    struct pi_batch* batch = nullptr;   /* the batch this model steps in, if any */
    hipStream_t own_stream = nullptr;   /* its stream while it is in a batch */
    Contrib con{};
    unsigned long long* d_count = nullptr;
    int32_t *dstamp = nullptr, *rstamp = nullptr;
    uint32_t* dcell = nullptr;
    float* rcons = nullptr;
    /* binning: new cell per store entry, stayer flags and their prefix, the
     * previous step's segments, queued entries, per-cell queued counts,
     * re-sorted cells; per-slot flags for the division marks */
    uint32_t *newc = nullptr, *stayf = nullptr, *P = nullptr, *segp = nullptr, *small = nullptr;
    uint32_t *cnts = nullptr, *mix = nullptr;
    uint8_t* sflag = nullptr;
    bool dirty = false;
    long long host_iter = 0;
    float last_dt = 0.0f;
    std::vector<void*> allocs;
    bool timing = false;
    std::vector<KTimer> timers;
    /* one steady step (vel₀ = vel₁, no timing) captured as a HIP graph, replayed
     * by pi_run; dropped when a launch argument changes (velocity buffers,
     * diagnostic tracers, Δt) */
    hipGraphExec_t gexec = nullptr;
    float gdt = 0.0f;
    size_t tcount = 0;
    long long shape[PI_NSHAPE] = {16, 16, 16384, 4096, 8, 1 << 22};   /* pi_set_launch_shape */
    int mode = PI_MODE_QUOTA;  /* physiology (pi_model_desc::mode) */
    float* dbd = nullptr;      /* CarbonMode: per slot, the Bd of an occupant dying in the step */
    /* flux boundary conditions of the tracers (pi_set_tracer_bc): the table on
     * the host and on the device, each face array's own buffer, the nt of each
     * time-dependent one */
    BcTab h_bc{};
    BcTab* d_bc = nullptr;
    float* bc_buf[PI_NTRACERS][PI_NBC]{};
    long long bc_nt[PI_NTRACERS][PI_NBC]{};
    /* abiotic particles: nsa species of cap slots, their state, the scratch of
     * particles_from_bcs! (cell flags, free slots, prefixes), num, and their
     * flux boundary conditions */
    int nsa = 0;
    float aparams[PI_MAXSA][PI_NAPARAMS]{};
    float *ax = nullptr, *ay = nullptr, *az = nullptr, *anum = nullptr;
    int32_t *axi = nullptr, *ayi = nullptr, *azi = nullptr;
    uint8_t* aac = nullptr;
    uint32_t *acflag = nullptr, *acpre = nullptr, *afflag = nullptr, *afpre = nullptr, *aflist = nullptr,
             *apart = nullptr;
    AbioBc h_abc{};
    AbioBc* d_abc = nullptr;
    float* abc_buf[PI_MAXSA][PI_NBC]{};
    long long abc_nt[PI_MAXSA][PI_NBC]{};
    /* phytoplankton-particle interaction and release: the pairs, the scratch of
     * the interaction kernels, the per-slot index of a reborn daughter, diags.phytos.spN.ptc */
    int npint = 0, nprel = 0, mc = 25;
    int pint[PI_MAXPAIRS][2]{}, prel[PI_MAXPAIRS][2]{};
    uint32_t *icnt = nullptr, *ioff = nullptr, *ifill = nullptr, *ilist = nullptr, *rflag = nullptr,
             *rrank = nullptr, *ifflag = nullptr, *ifpre = nullptr, *iflist = nullptr, *rdidx = nullptr;
    float* diag_ptc = nullptr;
};

static hipError_t dalloc(pi_model* m, void** p, size_t bytes)
{
    const size_t rb = ((bytes ? bytes : 4) + 255) & ~size_t(255);
    m->bytes += (long long)rb;
    if (m->dry) {                  /* footprint only: a placeholder address */
        *p = (void*)(uintptr_t)0x100;
        return hipSuccess;
    }
    hipError_t e = hipMalloc(p, bytes ? bytes : 4);
    if (e == hipSuccess) {
        m->allocs.push_back(*p);
        e = hipMemsetAsync(*p, 0, bytes ? bytes : 4, m->stream);   /* on the model's stream: no null-stream ordering */
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);  /* before any copy into it */
    }
    return e;
}
#define DALLOC(ptr, bytes) HIPCHECK(dalloc(m, (void**)&(ptr), (size_t)(bytes)))
/* a device call of create_impl, skipped when only the footprint is computed */
#define DEVCALL(x) do { if (!m->dry) HIPCHECK(x); } while (0)

/* ------------------------------------------------------------------ timing */
static void tbegin(pi_model* m, const char* name)
{
    if (!m->timing) return;
    if (m->tcount == m->timers.size()) {
        KTimer t{name, nullptr, nullptr};
        (void)hipEventCreate(&t.a);
        (void)hipEventCreate(&t.b);
        m->timers.push_back(t);
    }
    m->timers[m->tcount].name = name;
    (void)hipEventRecord(m->timers[m->tcount].a, m->stream);
}
static void tend(pi_model* m)
{
    if (!m->timing) return;
    (void)hipEventRecord(m->timers[m->tcount].b, m->stream);
    m->tcount++;
}

/* ------------------------------------------------------------------ helpers */
static int blocks_for(long long n) { return grid_for(n, 256, 4096); }

static pi_status sync_state(pi_model* m)
{
    HIPCHECK(hipStreamSynchronize(m->stream));
    HIPCHECK(hipMemcpy(&m->h_state, m->d_state, sizeof(DevState), hipMemcpyDeviceToHost));
    if (m->h_state.err == PI_ERR_HIP) return fail(PI_ERR_HIP, "device-side consistency check failed in division");
    if (m->h_state.err == PI_ERR_CAPACITY) {
        char buf[200];
        const double day = (double)(m->h_state.iteration + 1) * (double)m->last_dt / 86400.0;
        if (m->h_state.err_sp >= PI_ERR_SP_RELEASE)   /* particle_release! (particle_interaction.jl:131-134) */
            snprintf(buf, sizeof buf, "number of abiotic particles exceeds the capacity at timestep %g days "
                     "(release into abiotic species %u)", day, m->h_state.err_sp - PI_ERR_SP_RELEASE + 1);
        else if (m->h_state.err_sp >= PI_ERR_SP_ABIO)      /* particles_from_bcs! (particle_interaction.jl:256-259) */
            snprintf(buf, sizeof buf, "number of abiotic particles exceeds the capacity at timestep %g days "
                     "(abiotic species %u)", day, m->h_state.err_sp - PI_ERR_SP_ABIO + 1);
        else
            snprintf(buf, sizeof buf, "number of individual exceeds the capacity at timestep %g days (species %u)",
                     day, m->h_state.err_sp + 1);
        return fail(PI_ERR_CAPACITY, buf);
    }
    return PI_OK;
}

static pi_status rebuild_if_dirty(pi_model* m)
{
    if (!m->dirty) return PI_OK;
    static const uint32_t zero[2] = {0, 0};
    static_assert(offsetof(DevState, nA) == offsetof(DevState, nstore) + 4, "nstore, nA adjacent");
    HIPCHECK(hipMemcpyAsync((char*)m->d_state + offsetof(DevState, nstore), zero, sizeof zero,
                            hipMemcpyHostToDevice, m->stream));   /* nstore = nA = 0: all queued */
    launch_rebuild(m->stream, m->d_state, m->tab, m->cap, m->nsp, m->sflag);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(m->stream));
    m->dirty = false;
    return PI_OK;
}

/* host-side fill_halo_* of an input velocity (halo_regions.jl:64-99) */
static void host_fill(const pi_model* m, std::vector<float>& c, int dim, int mode)
{
    /* mode 0: lo, 1: hi, 2: hi_vel */
    const DevGrid& g = m->g;
    const int N = dim == 0 ? g.Nx : dim == 1 ? g.Ny : g.Nz;
    const int H = dim == 0 ? g.Hx : dim == 1 ? g.Hy : g.Hz;
    const int topo = dim == 0 ? g.tx : dim == 1 ? g.ty : g.tz;
    const long long stride = dim == 0 ? 1 : dim == 1 ? g.Px : (long long)g.Px * g.Py;
    const int P1 = dim == 0 ? g.Py : g.Px, P2 = dim == 2 ? g.Py : g.Pz;
    for (int o2 = 1; o2 <= P2; ++o2)
        for (int o1 = 1; o1 <= P1; ++o1) {
            long long base;
            if (dim == 0) base = (long long)(o1 - 1) * g.Px + (long long)(o2 - 1) * g.Px * g.Py;
            else if (dim == 1) base = (long long)(o1 - 1) + (long long)(o2 - 1) * g.Px * g.Py;
            else base = (long long)(o1 - 1) + (long long)(o2 - 1) * g.Px;
            auto at = [&](int p) -> float& { return c[base + (long long)(p - 1) * stride]; };
            float tmp[16];
            if (mode == 0) {
                if (topo == PI_PERIODIC) { for (int h = 0; h < H; ++h) tmp[h] = at(N + 1 + h); for (int h = 0; h < H; ++h) at(1 + h) = tmp[h]; }
                else { float v = at(H + 1); for (int h = 0; h < H; ++h) at(1 + h) = v; }
            } else if (mode == 1 || topo == PI_PERIODIC) {
                if (topo == PI_PERIODIC) { for (int h = 0; h < H; ++h) tmp[h] = at(H + 1 + h); for (int h = 0; h < H; ++h) at(N + H + 1 + h) = tmp[h]; }
                else { float v = at(N + H); for (int h = 0; h < H; ++h) at(N + H + 1 + h) = v; }
            } else {   /* hi_vel, bounded: c[N+H+2:N+2H] = c[N+H+1] */
                float v = at(N + H + 1);
                for (int h = 1; h < H; ++h) at(N + H + 1 + h) = v;
            }
        }
}

static pi_status create_impl(const pi_model_desc* d, pi_model* m)
{
    const pi_grid_desc& gd = d->grid;
    if (!m->dry) {
        HIPCHECK(hipGetDevice(&m->dev));
        HIPCHECK(hipDeviceGetAttribute(&m->ncu, hipDeviceAttributeMultiprocessorCount, m->dev));
        HIPCHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        HIPCHECK(hipStreamCreateWithFlags(&m->side, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming));
        HIPCHECK(hipEventCreateWithFlags(&m->ev_join, hipEventDisableTiming));
    }
    DevGrid& g = m->g;
    g.Nx = gd.Nx; g.Ny = gd.Ny; g.Nz = gd.Nz; g.Hx = gd.Hx; g.Hy = gd.Hy; g.Hz = gd.Hz;
    g.Px = gd.Nx + 2 * gd.Hx; g.Py = gd.Ny + 2 * gd.Hy; g.Pz = gd.Nz + 2 * gd.Hz;
    g.tx = gd.topo_x; g.ty = gd.topo_y; g.tz = gd.topo_z;
    g.dx = gd.dx; g.dy = gd.dy;
    g.ncell = (long long)g.Px * g.Py * g.Pz;
    if (g.ncell >= (1ll << 30)) return fail(PI_ERR_ARG, "grid too large: the padded grid must have fewer than 2^30 cells");
    m->ncell = g.ncell;
    m->nint = (long long)g.Nx * g.Ny * g.Nz;
    m->nsp = d->n_species;
    m->cap = d->capacity;
    m->ctot = m->cap * m->nsp;
    if (m->ctot >= (1ll << 32)) return fail(PI_ERR_ARG, "total capacity too large");
    m->nw = (int)((m->cap + 31) / 32);
    m->seed = d->rng_seed;
    m->h_params.assign(d->params, d->params + m->nsp * PI_NPARAMS);
    m->bgc.assign(d->bgc, d->bgc + PI_NBGC);
    float *dzC, *dzF;
    DALLOC(dzC, g.Pz * 4); DALLOC(dzF, g.Pz * 4);
    DEVCALL(hipMemcpy(dzC, gd.dzC, g.Pz * 4, hipMemcpyHostToDevice));
    DEVCALL(hipMemcpy(dzF, gd.dzF, g.Pz * 4, hipMemcpyHostToDevice));
    g.dzC = dzC; g.dzF = dzF;
    {   /* RN_double(1/b) of the fixed divisors Δx, Δy, 6, dzC[k], dzF[k] and of the
         * cell volume (Δx·Δy)·dzF[k] (divc in pi_device.h) */
        const int Pz = g.Pz;
        g.ydx = 1.0 / (double)g.dx; g.ydy = 1.0 / (double)g.dy; g.y6 = 1.0 / 6.0;
        std::vector<double> t(3 * (size_t)Pz);
        std::vector<float> vol(Pz);
        for (int k = 0; k < Pz; ++k) {
            vol[k] = (g.dx * g.dy) * gd.dzF[k];
            t[k] = 1.0 / (double)gd.dzC[k];
            t[Pz + k] = 1.0 / (double)gd.dzF[k];
            t[2 * Pz + k] = 1.0 / (double)vol[k];
        }
        double* tabs;
        float* dvol;
        DALLOC(tabs, 3ll * Pz * 8);
        DALLOC(dvol, 4ll * Pz);
        DEVCALL(hipMemcpy(tabs, t.data(), 3ll * Pz * 8, hipMemcpyHostToDevice));
        DEVCALL(hipMemcpy(dvol, vol.data(), 4ll * Pz, hipMemcpyHostToDevice));
        g.ydzC = tabs; g.ydzF = tabs + Pz; g.yvol = tabs + 2 * Pz; g.vol = dvol;
    }
    g.gm = 0;
    if (d->metrics) {      /* LatLonGrid: the metric arrays, with halos */
        const pi_metrics& mt = *d->metrics;
        const float* src[8] = {mt.dxC, mt.dyC, mt.dxF, mt.dyF, mt.Az, mt.Ax, mt.Ay, mt.Vol};
        const float** dst[8] = {&g.mdxC, &g.mdyC, &g.mdxF, &g.mdyF, &g.mAz, &g.mAx, &g.mAy, &g.mVol};
        for (int a = 0; a < 8; ++a) {
            const long long n = a < 5 ? (long long)g.Px * g.Py : g.ncell;
            float* p;
            DALLOC(p, n * 4);
            DEVCALL(hipMemcpy(p, src[a], n * 4, hipMemcpyHostToDevice));
            *dst[a] = p;
        }
        g.gm = 1;
    }
    g.has_land = 0; g.lm = nullptr;
    if (gd.landmask) {
        bool any0 = false;
        for (long long c = 0; c < g.ncell; ++c) if (!gd.landmask[c]) { any0 = true; break; }
        if (any0) {
            uint8_t* lm;
            DALLOC(lm, g.ncell);
            DEVCALL(hipMemcpy(lm, gd.landmask, g.ncell, hipMemcpyHostToDevice));
            g.lm = lm; g.has_land = 1;
        }
    }
    DALLOC(m->d_params, m->h_params.size() * 4);
    DEVCALL(hipMemcpy(m->d_params, m->h_params.data(), m->h_params.size() * 4, hipMemcpyHostToDevice));
    DALLOC(m->d_bc, sizeof(BcTab));    /* zeroed: no flux boundary conditions */
    m->nsa = d->n_abiotic;
    if (m->nsa > 0) {                  /* abiotic particles: state and scratch, all zeroed */
        const long long ns = (long long)m->nsa * m->cap, nc = (long long)m->nsa * g.ncell;
        for (int sa = 0; sa < m->nsa; ++sa)
            for (int q = 0; q < PI_NAPARAMS; ++q) m->aparams[sa][q] = d->abiotic_params[sa * PI_NAPARAMS + q];
        DALLOC(m->ax, ns * 4); DALLOC(m->ay, ns * 4); DALLOC(m->az, ns * 4);
        DALLOC(m->axi, ns * 4); DALLOC(m->ayi, ns * 4); DALLOC(m->azi, ns * 4);
        DALLOC(m->aac, ns);
        DALLOC(m->acflag, nc * 4); DALLOC(m->acpre, (nc + 1) * 4);
        DALLOC(m->afflag, ns * 4); DALLOC(m->afpre, (ns + 1) * 4); DALLOC(m->aflist, ns * 4);
        DALLOC(m->apart, (long long)scan_tiles((uint32_t)std::max(nc, ns)) * 4 + 64);
        DALLOC(m->anum, nc * 4);
        /* diags.phytos.spN.ptc (IronEnergyMode: the only mode whose ptc can change) */
        if (d->mode == PI_MODE_IRONENERGY) DALLOC(m->diag_ptc, (long long)m->nsp * g.ncell * 4);
        m->npint = d->n_interactions; m->nprel = d->n_releases; m->mc = d->max_candidates;
        for (int q = 0; q < m->npint; ++q) { m->pint[q][0] = d->interactions[2 * q]; m->pint[q][1] = d->interactions[2 * q + 1]; }
        for (int q = 0; q < m->nprel; ++q) { m->prel[q][0] = d->releases[2 * q]; m->prel[q][1] = d->releases[2 * q + 1]; }
        if (m->npint + m->nprel > 0) {
            /* scan outputs hold the total after the last entry */
            DALLOC(m->icnt, g.ncell * 4); DALLOC(m->ioff, (g.ncell + 1) * 4); DALLOC(m->ifill, g.ncell * 4);
            DALLOC(m->ilist, m->cap * 4); DALLOC(m->rflag, m->cap * 4); DALLOC(m->rrank, (m->cap + 1) * 4);
            DALLOC(m->ifflag, m->cap * 4); DALLOC(m->ifpre, (m->cap + 1) * 4); DALLOC(m->iflist, m->cap * 4);
            DALLOC(m->rdidx, m->ctot * 4);
        }
        DALLOC(m->d_abc, sizeof(AbioBc));
    }
    /* store x2 */
    DevState& hs = m->h_state;
    memset(&hs, 0, sizeof hs);
    /* a step's output holds the live individuals of its input (those dying
     * this step stay as inactive entries) plus the daughters: <= 2 x capacity */
    const long long scap = 2 * m->ctot;
    const int nex = mode_nex(d->mode);   /* extra state columns of the mode */
    for (int b = 0; b < 2; ++b) {
        Store& S = hs.store[b];
        DALLOC(S.key, scap * 4);
        for (float** f : {&S.x, &S.y, &S.z, &S.Sz, &S.Bm, &S.Cq, &S.Nq, &S.Pq, &S.Chl, &S.gen, &S.age})
            DALLOC(*f, scap * 4);
        for (int k = 0; k < nex; ++k) DALLOC(S.ex[k], scap * 4);
        DALLOC(S.flags, scap);
        DALLOC(S.cell, scap * 4);
    }
    /* slot table */
    SlotTable& T = m->tab;
    for (float** f : {&T.x, &T.y, &T.z, &T.Sz, &T.Bm, &T.Cq, &T.Nq, &T.Pq, &T.Chl, &T.gen, &T.age})
        DALLOC(*f, m->ctot * 4);
    for (int k = 0; k < nex; ++k) DALLOC(T.ex[k], m->ctot * 4);
    DALLOC(T.flags, m->ctot);
    DALLOC(T.cell, m->ctot * 4);
    DALLOC(T.death, m->ctot * 4);
    /* grid fields */
    const long long nc = g.ncell;
    /* tracer sets as slabs of 14 fields (field q at slab + q·ncell) */
    for (int b = 0; b < 2; ++b) {
        DALLOC(hs.tr[b][0], PI_NTRACERS * nc * 4);
        for (int q = 1; q < PI_NTRACERS; ++q) hs.tr[b][q] = hs.tr[b][0] + q * nc;
    }
    DALLOC(m->T1[0], PI_NTRACERS * nc * 4); DALLOC(m->T2[0], PI_NTRACERS * nc * 4);
    for (int q = 1; q < PI_NTRACERS; ++q) { m->T1[q] = m->T1[0] + q * nc; m->T2[q] = m->T2[0] + q * nc; }
    for (int b = 0; b < 3; ++b) for (int c = 0; c < 3; ++c) DALLOC(m->vel[b][c], nc * 4);
    m->v0 = 0; m->v1 = 0; m->vel_same = true;
    DALLOC(hs.par[0], nc * 4); DALLOC(hs.par[1], nc * 4);
    DALLOC(m->Chl, nc * 4); DALLOC(m->pop, nc * 4); DALLOC(m->temp, nc * 4);
    DALLOC(m->PARF_in, (long long)g.Nx * g.Ny * 4); DALLOC(m->PARF_out, (long long)g.Nx * g.Ny * 4);
    const int plk_ids[10] = {PI_DIC, PI_NH4, PI_NO3, PI_PO4, PI_DOC, PI_DON, PI_DOP, PI_POC, PI_PON, PI_POP};
    DALLOC(m->plk_slab, PI_NTRACERS * nc * 4);
    for (int q : plk_ids) m->plk[q] = m->plk_slab + q * nc;
    if (d->mode == PI_MODE_IRONENERGY)   /* its losses and uptake reach DFe and PFe_bio too */
        for (int q : {PI_DFe, PI_PFe_bio}) m->plk[q] = m->plk_slab + q * nc;
    DALLOC(m->diagPAR, nc * 4);
    DALLOC(m->diag, (long long)m->nsp * 4 * nc * 4);
    /* sort / scan scratch */
    DALLOC(m->cnt, nc * 4);
    DALLOC(m->seg, (nc + 1) * 4);
    const long long scan_n = std::max<long long>(std::max<long long>(nc, 2ll * m->nsp * m->nw), 2 * m->ctot);
    DALLOC(m->partials, ((scan_n + 2047) / 2048 + 1) * 4);
    DALLOC(m->pkey, m->ctot * 4); DALLOC(m->psrc, m->ctot * 4);
    DALLOC(m->skey, 2 * m->ctot * 4); DALLOC(m->ssrc, 2 * m->ctot * 4);
    DALLOC(m->big, nc * 4);
    DALLOC(m->lng, nc * 4);
    m->long_blocks = (int)std::min<long long>((long long)g.Nx * g.Ny * g.Nz, 512);
    DALLOC(m->bm, 2ll * m->nsp * m->nw * 4);
    DALLOC(m->G, (2ll * m->nsp * m->nw + 1) * 4);
    for (int q = 0; q < mode_ncon(d->mode); ++q) DALLOC(m->con.v[q], m->ctot * 4);
    DALLOC(m->con.tag, m->ctot);
    DALLOC(m->d_count, 8);
    DALLOC(m->newc, 2 * m->ctot * 4); DALLOC(m->stayf, 2 * m->ctot * 4); DALLOC(m->P, (2 * m->ctot + 1) * 4);
    DALLOC(m->segp, (nc + 1) * 4);
    /* queue shards: shard q holds the entries e < 2·ctot with (e / 256) mod PI_NQ = q */
    m->qcap = (uint32_t)(((2 * m->ctot + 255) / 256 + PI_NQ - 1) / PI_NQ * 256);
    DALLOC(m->small, (long long)m->qcap * PI_NQ * 4);
    DALLOC(m->cnts, nc * 4);
    DALLOC(m->mix, (long long)m->qcap * PI_NQ * 4);   /* shards as the queue's (by queue position) */
    DALLOC(m->sflag, m->ctot);
    DALLOC(m->dstamp, m->ctot * 4); DALLOC(m->rstamp, m->ctot * 4);
    DALLOC(m->dcell, m->ctot * 4); DALLOC(m->rcons, m->ctot * PI_RCONS * 4);
    m->mode = d->mode;
    if (m->mode == PI_MODE_CARBON || m->mode == PI_MODE_IRONENERGY) DALLOC(m->dbd, m->ctot * 4);
    if (!m->dry) {   /* stamps start below any iteration */
        std::vector<int32_t> neg(m->ctot, -1);
        DEVCALL(hipMemcpy(m->dstamp, neg.data(), m->ctot * 4, hipMemcpyHostToDevice));
        DEVCALL(hipMemcpy(m->rstamp, neg.data(), m->ctot * 4, hipMemcpyHostToDevice));
    }
    DALLOC(m->d_state, sizeof(DevState));
    hs.scur = hs.tcur = hs.pcur = 0;
    hs.iteration = 0;
    hs.nstore = 0;
    DEVCALL(hipMemcpy(m->d_state, &hs, sizeof hs, hipMemcpyHostToDevice));
    for (int q = 0; q < d->n_diag_tracers; ++q) {
        if (d->diag_tracers[q] < 0 || d->diag_tracers[q] >= PI_NTRACERS) return fail(PI_ERR_ARG, "diag tracer id");
        float* p; DALLOC(p, nc * 4);
        m->diag_ids.push_back(d->diag_tracers[q]); m->diag_tr.push_back(p);
    }
    return PI_OK;
}

/* ======================================================================= ABI */
extern "C" {

const char* pi_last_error(void) { return g_err.c_str(); }
const char* pi_version(void) { return "libpi_hip 0.1 (gfx950)"; }

pi_status pi_init(int32_t device)
{
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(PI_ERR_ARG, "no such HIP device");
    HIPCHECK(hipSetDevice(device));
    char* base = nullptr;               /* the static scratch of pi_batch_run_noalloc on this device */
    return scratch_base(device, &base);
}

static pi_status check_desc(const pi_model_desc* d)
{
    const pi_grid_desc& gd = d->grid;
    if (gd.Nx < 1 || gd.Ny < 1 || gd.Nz < 1 || gd.Hx < 2 || gd.Hy < 2 || gd.Hz < 2 || gd.Hx > 8 || gd.Hy > 8 || gd.Hz > 8)
        return fail(PI_ERR_ARG, "grid: N >= 1 and 2 <= halo <= 8 required");
    if (gd.topo_z != PI_BOUNDED) return fail(PI_ERR_ARG, "grid: z topology must be Bounded");
    if (d->n_species < 1 || d->n_species > PI_MAXSP) return fail(PI_ERR_ARG, "n_species out of range");
    if (d->capacity < 1 || d->capacity > (long long)PI_SLOT_MASK) return fail(PI_ERR_ARG, "capacity out of range");
    if (d->mode < PI_MODE_QUOTA || d->mode > PI_MODE_IRONENERGY) return fail(PI_ERR_UNSUPPORTED, "physiology mode");
    if (d->metrics) {
        const pi_metrics& mt = *d->metrics;
        if (!mt.dxC || !mt.dyC || !mt.dxF || !mt.dyF || !mt.Az || !mt.Ax || !mt.Ay || !mt.Vol)
            return fail(PI_ERR_ARG, "grid metrics: every array is required");
    }
    if (d->n_abiotic < 0 || d->n_abiotic > PI_MAXSA) return fail(PI_ERR_ARG, "n_abiotic out of range [0, 4]");
    if (d->n_abiotic > 0 && !d->abiotic_params) return fail(PI_ERR_ARG, "abiotic particle parameters missing");
    if (d->n_interactions < 0 || d->n_interactions > PI_MAXPAIRS || d->n_releases < 0 || d->n_releases > PI_MAXPAIRS)
        return fail(PI_ERR_UNSUPPORTED, "at most 4 interaction and 4 release pairs");
    if ((d->n_interactions && !d->interactions) || (d->n_releases && !d->releases))
        return fail(PI_ERR_ARG, "interaction / release pairs missing");
    if (d->n_interactions && d->mode != PI_MODE_IRONENERGY)
        return fail(PI_ERR_ARG, "phytoplankton-particle interaction needs IronEnergyMode (max_ptc, ptc)");
    if (d->n_interactions + d->n_releases > 0 && (d->max_candidates < 1 || d->max_candidates > PI_MAXCAND))
        return fail(PI_ERR_UNSUPPORTED, "max_candidates must be in [1, 64]");
    for (int q = 0; q < d->n_interactions + d->n_releases; ++q) {
        const int32_t* pr = q < d->n_interactions ? d->interactions + 2 * q : d->releases + 2 * (q - d->n_interactions);
        if (pr[0] < 0 || pr[0] >= d->n_species || pr[1] < 0 || pr[1] >= d->n_abiotic)
            return fail(PI_ERR_ARG, "interaction / release pair out of range");
    }
    {
        const long long nc = (long long)(gd.Nx + 2 * gd.Hx) * (gd.Ny + 2 * gd.Hy) * (gd.Nz + 2 * gd.Hz);
        if ((long long)d->n_abiotic * nc >= (1ll << 31) || (long long)d->n_abiotic * d->capacity >= (1ll << 31))
            return fail(PI_ERR_ARG, "abiotic particles: species x cells and species x capacity must stay below 2^31");
    }
    const int dvt = d->mode == PI_MODE_CARBON ? (int)PI_C_dvid_type
                  : d->mode == PI_MODE_IRONENERGY ? (int)PI_I_dvid_type : (int)PI_P_dvid_type;
    for (int s = 0; s < d->n_species && d->mode != PI_MODE_MACROMOLECULAR; ++s) {   /* MM: no dvid_type */
        float t = d->params[s * PI_NPARAMS + dvt];
        if (!(t == 1.0f || t == 2.0f || t == 3.0f)) return fail(PI_ERR_ARG, "Wrong cell division type, must be in 1 to 3");
    }
    return PI_OK;
}

pi_status pi_footprint(const pi_model_desc* d, int64_t* bytes)
{
    if (!d || !bytes) return fail(PI_ERR_ARG, "null argument");
    pi_status st = check_desc(d);
    if (st != PI_OK) return st;
    pi_model m;
    m.dry = true;
    st = create_impl(d, &m);
    if (st != PI_OK) return st;
    *bytes = m.bytes;
    return PI_OK;
}

pi_status pi_model_footprint(const pi_model* m, int64_t* bytes)
{
    if (!m || !bytes) return fail(PI_ERR_ARG, "null argument");
    *bytes = m->bytes;
    return PI_OK;
}

pi_status pi_model_create(const pi_model_desc* d, pi_model** out)
{
    if (!d || !out) return fail(PI_ERR_ARG, "null argument");
    pi_status cs = check_desc(d);
    if (cs != PI_OK) return cs;
    pi_model* m = new pi_model();
    pi_status st = create_impl(d, m);
    if (st != PI_OK) {
        for (void* p : m->allocs) (void)hipFree(p);
        if (m->stream) (void)hipStreamDestroy(m->stream);
        if (m->side) (void)hipStreamDestroy(m->side);
        if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
        if (m->ev_join) (void)hipEventDestroy(m->ev_join);
        delete m;
        return st;
    }
    *out = m;
    return PI_OK;
}

static void drop_graph(pi_model* m)
{
    if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
    m->gexec = nullptr;
}

pi_status pi_model_destroy(pi_model* m)
{
    if (!m) return PI_OK;
    if (m->batch) (void)pi_batch_remove(m->batch, m);
    (void)hipStreamSynchronize(m->stream);
    drop_graph(m);
    for (auto& t : m->timers) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (void* p : m->allocs) (void)hipFree(p);
    for (auto& row : m->bc_buf) for (float* p : row) if (p) (void)hipFree(p);
    for (auto& row : m->abc_buf) for (float* p : row) if (p) (void)hipFree(p);
    (void)hipStreamDestroy(m->stream);
    (void)hipStreamDestroy(m->side);
    (void)hipEventDestroy(m->ev_fork);
    (void)hipEventDestroy(m->ev_join);
    delete m;
    return PI_OK;
}

pi_status pi_set_launch_shape(pi_model* m, int32_t which, int64_t value)
{
    static const long long lo[PI_NSHAPE] = {1, 1, 1, 1, 1, 1};
    static const long long hi[PI_NSHAPE] = {1024, 4096, 1 << 24, 1 << 24, 64, 1 << 24};
    if (!m) return fail(PI_ERR_ARG, "null model");
    if (which < 0 || which >= PI_NSHAPE) return fail(PI_ERR_ARG, "launch shape: unknown parameter");
    if (value < lo[which] || value > hi[which]) return fail(PI_ERR_ARG, "launch shape: value out of range");
    if (m->shape[which] == value) return PI_OK;
    HIPCHECK(hipStreamSynchronize(m->stream));      /* a replayed step may still run */
    drop_graph(m);
    m->shape[which] = value;
    return PI_OK;
}

pi_status pi_get_launch_shape(const pi_model* m, int32_t which, int64_t* value)
{
    if (!m || !value) return fail(PI_ERR_ARG, "null argument");
    if (which < 0 || which >= PI_NSHAPE) return fail(PI_ERR_ARG, "launch shape: unknown parameter");
    *value = m->shape[which];
    return PI_OK;
}

pi_status pi_set_diag_tracers(pi_model* m, int32_t n, const int32_t* ids)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    HIPCHECK(hipStreamSynchronize(m->stream));
    for (int32_t a = 0; a < n; ++a) if (ids[a] < 0 || ids[a] >= PI_NTRACERS) return fail(PI_ERR_ARG, "tracer id");
    /* reuse or allocate accumulators */
    drop_graph(m);
    std::vector<float*> keep = m->diag_tr;
    m->diag_tr.clear(); m->diag_ids.clear();
    for (int32_t a = 0; a < n; ++a) {
        float* p;
        if (!keep.empty()) { p = keep.back(); keep.pop_back(); HIPCHECK(hipMemset(p, 0, m->ncell * 4)); }
        else DALLOC(p, m->ncell * 4);
        m->diag_ids.push_back(ids[a]); m->diag_tr.push_back(p);
    }
    return PI_OK;
}

pi_status pi_upload_individuals(pi_model* m, int32_t sp, const pi_individuals_host* h)
{
    if (!m || !h) return fail(PI_ERR_ARG, "null argument");
    if (sp < 0 || sp >= m->nsp) return fail(PI_ERR_ARG, "species out of range");
    pi_status s = rebuild_if_dirty(m);
    if (s != PI_OK) return s;
    s = sync_state(m);
    if (s != PI_OK) return s;
    /* keep the other species' live state: flush the store into the table */
    launch_store_to_table(m->stream, m->d_state, m->tab, m->cap, 2 * m->ctot);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(m->stream));
    const long long cap = m->cap, off = (long long)sp * cap;
    const DevGrid& g = m->g;
    std::vector<uint8_t> flags(cap);
    std::vector<uint32_t> cell(cap);
    std::vector<int32_t> death(cap, (int32_t)m->h_state.iteration);
    for (long long i = 0; i < cap; ++i) {
        uint8_t f = (h->ac && h->ac[i]) ? F_AC : 0;
        const float* dec[3] = {h->graz, h->mort, h->dvid};
        const uint8_t bits[3] = {F_GRAZ, F_MORT, F_DVID};
        for (int k = 0; k < 3; ++k) {
            float v = dec[k] ? dec[k][i] : 0.0f;
            if (v != 0.0f && v != 1.0f) return fail(PI_ERR_ARG, "graz/mort/dvid decisions must be 0 or 1");
            if (v == 1.0f) f |= bits[k];
        }
        long long xi = h->xi ? h->xi[i] : 0, yi = h->yi ? h->yi[i] : 0, zi = h->zi ? h->zi[i] : 0;
        if (xi == 0 && yi == 0 && zi == 0) { f |= F_ZERO; cell[i] = 0; }
        else {
            if (xi < 1 || xi > g.Px || yi < 1 || yi > g.Py || zi < 1 || zi > g.Pz)
                return fail(PI_ERR_ARG, "xi/yi/zi outside the padded grid");
            cell[i] = (uint32_t)((xi - 1) + (long long)g.Px * ((yi - 1) + (long long)g.Py * (zi - 1)));
        }
        flags[i] = f;
    }
    const SlotTable& T = m->tab;
    const float* src[11 + PI_NEX] = {h->x, h->y, h->z, h->Sz, h->Bm, h->Cq, h->Nq, h->Pq, h->Chl, h->gen, h->age};
    float* dst[11 + PI_NEX] = {T.x, T.y, T.z, T.Sz, T.Bm, T.Cq, T.Nq, T.Pq, T.Chl, T.gen, T.age};
    const int nex = mode_nex(m->mode);
    for (int k = 0; k < nex; ++k) { src[11 + k] = h->ex[k]; dst[11 + k] = T.ex[k]; }
    for (int k = 0; k < 11 + nex; ++k) {
        if (src[k]) HIPCHECK(hipMemcpy(dst[k] + off, src[k], cap * 4, hipMemcpyHostToDevice));
        else HIPCHECK(hipMemset(dst[k] + off, 0, cap * 4));
    }
    HIPCHECK(hipMemcpy(T.flags + off, flags.data(), cap, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(T.cell + off, cell.data(), cap * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(T.death + off, death.data(), cap * 4, hipMemcpyHostToDevice));
    m->dirty = true;
    return PI_OK;
}

/* The initial state generated on the device: every species' slots as
 * initialize_plankton! (QuotaMode/plankton_generation.jl:35-72) and the 14
 * tracers as generate_tracers with the (c0, rand_noise) source
 * (tracer_fields.jl:93-134) followed by fill_halo_tracer!, all drawn from the
 * shared stream: the values the host generator (ModelSpec) uploads. */
pi_status pi_generate_state(pi_model* m, const int64_t* n_individuals, const float* tracer_c0,
                            const float* tracer_lower, const int32_t* tracer_n)
{
    if (!m || !n_individuals || !tracer_c0 || !tracer_lower || !tracer_n) return fail(PI_ERR_ARG, "null argument");
    for (int s = 0; s < m->nsp; ++s)
        if (n_individuals[s] < 0 || n_individuals[s] > m->cap) return fail(PI_ERR_ARG, "individuals exceed the capacity");
    for (int q = 0; q < PI_NTRACERS; ++q)
        if (tracer_n[q] < 1 || !(tracer_c0[q] >= 0.0f)) return fail(PI_ERR_ARG, "tracer initial condition");
    pi_status st = rebuild_if_dirty(m);
    if (st != PI_OK) return st;
    st = sync_state(m);
    if (st != PI_OK) return st;
    InitArgs a{};
    a.g = m->g; a.st = m->d_state; a.tab = m->tab; a.params = m->d_params; a.cap = m->cap; a.nsp = m->nsp;
    for (int s = 0; s < m->nsp; ++s) a.N[s] = n_individuals[s];
    a.seed = m->seed;
    a.mode = m->mode;
    launch_init_individuals(m->stream, a);
    TracerInitArgs t{};
    t.g = m->g; t.slab = m->h_state.tr[m->h_state.tcur][0]; t.seed = m->seed; t.step = 1.0e-4f;
    for (int q = 0; q < PI_NTRACERS; ++q) { t.c0[q] = tracer_c0[q]; t.lower[q] = tracer_lower[q]; t.n[q] = tracer_n[q]; }
    launch_init_tracers(m->stream, t);
    const KInfo* K = kinfo();
    const DevGrid& g = m->g;
    FillArgs fa = zeroed<FillArgs>();
    fa.g = g; fa.st = m->d_state; fa.f = t.slab; fa.nf = PI_NTRACERS; fa.lines_full = 1;
    const long long lines[3] = {(long long)g.Py * g.Pz, (long long)g.Px * g.Pz, (long long)g.Px * g.Py};
    for (int d = 0; d < 3; ++d) {
        fa.dim = d;
        void* argp = &fa;
        HIPCHECK(hipLaunchKernel(K[K_FILL].one, dim3(blocks_for(lines[d] * PI_NTRACERS)), dim3(256), &argp, 0, m->stream));
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(m->stream));
    m->dirty = true;
    return PI_OK;
}

pi_status pi_upload_tracer(pi_model* m, int32_t q, const float* c)
{
    if (!m || !c || q < 0 || q >= PI_NTRACERS) return fail(PI_ERR_ARG, "bad tracer upload");
    pi_status s = sync_state(m);
    if (s != PI_OK) return s;
    HIPCHECK(hipMemcpy(m->h_state.tr[m->h_state.tcur][q], c, m->ncell * 4, hipMemcpyHostToDevice));
    return PI_OK;
}

pi_status pi_set_velocity(pi_model* m, const float* u, const float* v, const float* w)
{
    if (!m || !u || !v || !w) return fail(PI_ERR_ARG, "null velocity");
    const DevGrid& g = m->g;
    HIPCHECK(hipStreamSynchronize(m->stream));
    int b = 0;
    while (b == m->v0 || (b == m->v1 && m->v1 != m->v0)) ++b;   /* a set used by neither vel₀ nor vel₁ */
    if (b > 2) b = (m->v0 + 1) % 3;
    const int nx[3] = {g.Nx + (g.tx == PI_BOUNDED), g.Nx, g.Nx};
    const int ny[3] = {g.Ny, g.Ny + (g.ty == PI_BOUNDED), g.Ny};
    const int nz[3] = {g.Nz, g.Nz, g.Nz + (g.tz == PI_BOUNDED)};
    const float* in[3] = {u, v, w};
    std::vector<float> pad(m->ncell);
    for (int c = 0; c < 3; ++c) {
        std::fill(pad.begin(), pad.end(), 0.0f);
        HIPCHECK(hipMemcpy(pad.data(), m->vel[m->v1][c], m->ncell * 4, hipMemcpyDeviceToHost)); /* previous halos */
        for (int k = 0; k < nz[c]; ++k) for (int j = 0; j < ny[c]; ++j) for (int i = 0; i < nx[c]; ++i)
            pad[(long long)(g.Hx + i) + (long long)g.Px * ((g.Hy + j) + (long long)g.Py * (g.Hz + k))] =
                in[c][i + (long long)nx[c] * (j + (long long)ny[c] * k)];
        /* fill_halo_u!/v!/w! order */
        if (c == 0) { host_fill(m, pad, 0, 2); host_fill(m, pad, 0, 0); host_fill(m, pad, 1, 0); host_fill(m, pad, 1, 1); host_fill(m, pad, 2, 0); host_fill(m, pad, 2, 1); }
        if (c == 1) { host_fill(m, pad, 1, 2); host_fill(m, pad, 0, 0); host_fill(m, pad, 0, 1); host_fill(m, pad, 1, 0); host_fill(m, pad, 2, 0); host_fill(m, pad, 2, 1); }
        if (c == 2) { host_fill(m, pad, 2, 2); host_fill(m, pad, 0, 0); host_fill(m, pad, 0, 1); host_fill(m, pad, 1, 0); host_fill(m, pad, 1, 1); host_fill(m, pad, 2, 0); }
        HIPCHECK(hipMemcpy(m->vel[b][c], pad.data(), m->ncell * 4, hipMemcpyHostToDevice));
    }
    m->v1 = b;
    m->vel_same = false;
    drop_graph(m);
    return PI_OK;
}

pi_status pi_set_parf(pi_model* m, const float* parf)
{
    if (!m || !parf) return fail(PI_ERR_ARG, "null PARF");
    HIPCHECK(hipStreamSynchronize(m->stream));
    HIPCHECK(hipMemcpy(m->PARF_in, parf, (size_t)m->g.Nx * m->g.Ny * 4, hipMemcpyHostToDevice));
    return PI_OK;
}

pi_status pi_set_temperature(pi_model* m, const float* t)
{
    if (!m || !t) return fail(PI_ERR_ARG, "null temperature");
    const DevGrid& g = m->g;
    HIPCHECK(hipStreamSynchronize(m->stream));
    std::vector<float> pad(m->ncell, 0.0f);
    for (int k = 0; k < g.Nz; ++k) for (int j = 0; j < g.Ny; ++j) for (int i = 0; i < g.Nx; ++i)
        pad[(long long)(g.Hx + i) + (long long)g.Px * ((g.Hy + j) + (long long)g.Py * (g.Hz + k))] =
            t[i + (long long)g.Nx * (j + (long long)g.Ny * k)];
    HIPCHECK(hipMemcpy(m->temp, pad.data(), m->ncell * 4, hipMemcpyHostToDevice));
    return PI_OK;
}

pi_status pi_set_tracer_bc(pi_model* m, int32_t q, int32_t f, int32_t kind, const float* values, int64_t nt)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    if (q < 0 || q >= PI_NTRACERS) return fail(PI_ERR_ARG, "tracer id");
    if (f < 0 || f >= PI_NBC) return fail(PI_ERR_ARG, "boundary position");
    if (kind < PI_BCK_NONE || kind > PI_BCK_SERIES) return fail(PI_ERR_ARG, "boundary condition kind");
    if (kind != PI_BCK_NONE && !values) return fail(PI_ERR_ARG, "null boundary condition values");
    if (kind == PI_BCK_SERIES && nt < 1)
        return fail(PI_ERR_ARG, "BC: a time-dependent flux boundary condition needs nt >= 1");
    const DevGrid& g = m->g;
    const long long n1 = f <= PI_BC_EAST ? g.Ny : g.Nx, n2 = f <= PI_BC_NORTH ? g.Nz : g.Ny;
    const long long n = kind == PI_BCK_FIELD ? n1 * n2 : kind == PI_BCK_SERIES ? n1 * n2 * nt : 0;
    HIPCHECK(hipStreamSynchronize(m->stream));      /* a step may still read the old values */
    if (m->bc_buf[q][f]) {
        HIPCHECK(hipFree(m->bc_buf[q][f]));
        m->bc_buf[q][f] = nullptr;
    }
    m->h_bc.kind[q][f] = kind;
    m->h_bc.val[q][f] = kind == PI_BCK_CONST ? values[0] : 0.0f;
    m->h_bc.arr[q][f] = nullptr;
    m->bc_nt[q][f] = kind == PI_BCK_SERIES ? nt : 0;
    if (n) {
        float* d = nullptr;
        HIPCHECK(hipMalloc((void**)&d, (size_t)n * 4));
        m->bc_buf[q][f] = d;
        HIPCHECK(hipMemcpy(d, values, (size_t)n * 4, hipMemcpyHostToDevice));
        m->h_bc.arr[q][f] = d;
    }
    HIPCHECK(hipMemcpy(m->d_bc, &m->h_bc, sizeof(BcTab), hipMemcpyHostToDevice));
    drop_graph(m);
    return PI_OK;
}

/* A step reads a time-dependent flux boundary condition at the model's
 * iteration: refuse `n` more steps that would pass the end of one. */
static pi_status bc_steps_ok(const pi_model* m, int64_t n)
{
    for (int q = 0; q < PI_NTRACERS; ++q)
        for (int f = 0; f < PI_NBC; ++f)
            if (m->bc_nt[q][f] && m->host_iter + n > m->bc_nt[q][f])
                return fail(PI_ERR_ARG, "BC: the time-dependent flux boundary condition of a tracer has no value "
                                        "for this iteration (size(bc, 3) < iteration)");
    for (int sa = 0; sa < m->nsa; ++sa)
        for (int f = 0; f < PI_NBC; ++f)
            if (m->abc_nt[sa][f] && m->host_iter + n > m->abc_nt[sa][f])
                return fail(PI_ERR_ARG, "BC: the time-dependent flux boundary condition of an abiotic particle "
                                        "has no value for this iteration (size(bc, 3) < iteration)");
    return PI_OK;
}

pi_status pi_set_particle_bc(pi_model* m, int32_t sa, int32_t f, int32_t kind, const float* values, int64_t nt)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    if (sa < 0 || sa >= m->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
    if (f < 0 || f >= PI_NBC) return fail(PI_ERR_ARG, "boundary position");
    if (kind < PI_BCK_NONE || kind > PI_BCK_SERIES) return fail(PI_ERR_ARG, "boundary condition kind");
    if (kind != PI_BCK_NONE && !values) return fail(PI_ERR_ARG, "null boundary condition values");
    if (kind == PI_BCK_SERIES && nt < 1)
        return fail(PI_ERR_ARG, "BC: a time-dependent flux boundary condition needs nt >= 1");
    const DevGrid& g = m->g;
    const long long n1 = f <= PI_BC_EAST ? g.Ny : g.Nx, n2 = f <= PI_BC_NORTH ? g.Nz : g.Ny;
    const long long n = kind == PI_BCK_FIELD ? n1 * n2 : kind == PI_BCK_SERIES ? n1 * n2 * nt : 0;
    HIPCHECK(hipStreamSynchronize(m->stream));
    if (m->abc_buf[sa][f]) {
        HIPCHECK(hipFree(m->abc_buf[sa][f]));
        m->abc_buf[sa][f] = nullptr;
    }
    m->h_abc.kind[sa][f] = kind;
    m->h_abc.val[sa][f] = kind == PI_BCK_CONST ? values[0] : 0.0f;
    m->h_abc.arr[sa][f] = nullptr;
    m->abc_nt[sa][f] = kind == PI_BCK_SERIES ? nt : 0;
    if (n) {
        float* d = nullptr;
        HIPCHECK(hipMalloc((void**)&d, (size_t)n * 4));
        m->abc_buf[sa][f] = d;
        HIPCHECK(hipMemcpy(d, values, (size_t)n * 4, hipMemcpyHostToDevice));
        m->h_abc.arr[sa][f] = d;
    }
    HIPCHECK(hipMemcpy(m->d_abc, &m->h_abc, sizeof(AbioBc), hipMemcpyHostToDevice));
    drop_graph(m);
    return PI_OK;
}

pi_status pi_upload_abiotic(pi_model* m, int32_t sa, const float* x, const float* y, const float* z, const uint8_t* ac)
{
    if (!m || !x || !y || !z || !ac) return fail(PI_ERR_ARG, "null argument");
    if (sa < 0 || sa >= m->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
    HIPCHECK(hipStreamSynchronize(m->stream));
    const size_t o = (size_t)sa * m->cap, n = (size_t)m->cap;
    HIPCHECK(hipMemcpy(m->ax + o, x, n * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(m->ay + o, y, n * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(m->az + o, z, n * 4, hipMemcpyHostToDevice));
    std::vector<uint8_t> a(n);
    for (size_t i = 0; i < n; ++i) a[i] = ac[i] ? 1 : 0;
    HIPCHECK(hipMemcpy(m->aac + o, a.data(), n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(m->axi + o, 0, n * 4));       /* xi, yi, zi start at 0 */
    HIPCHECK(hipMemset(m->ayi + o, 0, n * 4));
    HIPCHECK(hipMemset(m->azi + o, 0, n * 4));
    return PI_OK;
}

pi_status pi_download_abiotic(pi_model* m, int32_t sa, int32_t field, void* host, int64_t n)
{
    if (!m || !host) return fail(PI_ERR_ARG, "null argument");
    if (sa < 0 || sa >= m->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
    if (field < PI_AB_x || field > PI_AB_ac) return fail(PI_ERR_ARG, "abiotic field id");
    if (n != m->cap) return fail(PI_ERR_ARG, "n must equal capacity");
    pi_status s = sync_state(m);
    if (s != PI_OK && s != PI_ERR_CAPACITY) return s;
    const size_t o = (size_t)sa * m->cap;
    if (field <= PI_AB_z) {
        const float* src = field == PI_AB_x ? m->ax : field == PI_AB_y ? m->ay : m->az;
        HIPCHECK(hipMemcpy(host, src + o, (size_t)n * 4, hipMemcpyDeviceToHost));
    } else if (field <= PI_AB_zi) {
        const int32_t* src = field == PI_AB_xi ? m->axi : field == PI_AB_yi ? m->ayi : m->azi;
        std::vector<int32_t> t((size_t)n);
        HIPCHECK(hipMemcpy(t.data(), src + o, (size_t)n * 4, hipMemcpyDeviceToHost));
        int64_t* out = (int64_t*)host;
        for (int64_t i = 0; i < n; ++i) out[i] = t[(size_t)i];
    } else {
        HIPCHECK(hipMemcpy(host, m->aac + o, (size_t)n, hipMemcpyDeviceToHost));
    }
    return PI_OK;
}

pi_status pi_download_abiotic_diag(pi_model* m, int32_t sa, float* h)
{
    if (!m || !h) return fail(PI_ERR_ARG, "null argument");
    if (sa < 0 || sa >= m->nsa) return fail(PI_ERR_ARG, "abiotic species out of range");
    HIPCHECK(hipStreamSynchronize(m->stream));
    HIPCHECK(hipMemcpy(h, m->anum + (size_t)sa * m->ncell, (size_t)m->ncell * 4, hipMemcpyDeviceToHost));
    return PI_OK;
}

/* ------------------------------------------------------------------ step */

static void plan_scan(Plan& P, const char* grp, const ScanIn& in, uint32_t* partials, uint32_t* out)
{
    ScanArgs a = zeroed<ScanArgs>();
    a.in = in; a.partials = partials; a.out = out; a.nb = scan_tiles(in.n);
    P.add(K_SCAN_P, grp, a.nb, a);
    P.add(K_SCAN_F, grp, a.nb, a);
}

static AbioArgs abio_args(const pi_model* m, float dt, const Vel& V0, const Vel& VH, const Vel& V1)
{
    AbioArgs a = zeroed<AbioArgs>();
    a.adv.g = m->g; a.adv.st = m->d_state; a.adv.v0 = V0; a.adv.vh = VH; a.adv.v1 = V1; a.adv.dt = dt;
    a.adv.sx = std::sqrt(m->bgc[PI_B_kappa_hP] * dt);
    a.adv.sy = std::sqrt(m->bgc[PI_B_kappa_hP] * dt);
    a.adv.sz = std::sqrt(m->bgc[PI_B_kappa_vP] * dt);
    a.adv.seed = m->seed;
    a.nsa = m->nsa; a.cap = m->cap; a.ncell = m->ncell;
    a.x = m->ax; a.y = m->ay; a.z = m->az; a.xi = m->axi; a.yi = m->ayi; a.zi = m->azi; a.ac = m->aac;
    a.cflag = m->acflag; a.cpre = m->acpre; a.fflag = m->afflag; a.fpre = m->afpre; a.flist = m->aflist;
    a.num = m->anum; a.bc = m->d_abc;
    a.num_late = m->npint + m->nprel > 0 ? 1 : 0;
    for (int sa = 0; sa < m->nsa; ++sa) {
        for (int f = 0; f < PI_NBC; ++f)
            if (m->h_abc.kind[sa][f] != PI_BCK_NONE) a.bcm |= 1u << (6 * sa + f);
        for (int q = 0; q < PI_NAPARAMS; ++q) a.p[sa][q] = m->aparams[sa][q];
    }
    return a;
}

/* Builds the plan of the model's next step and advances the host-side
 * bookkeeping of that step (vel₀ <- vel₁, iteration count). */
static void build_step(pi_model* m, float dt, Plan& P)
{
    const DevGrid& g = m->g;
    DevState* st = m->d_state;
    const long long pb = blocks_for(2 * m->ctot);
    const long long cb = blocks_for(m->ncell);
    P.clear();
    /* vel½ (time_step.jl:17-19); equal to vel₀ bit for bit when vel₀ == vel₁ */
    int vh = m->v0;
    {
        VelhalfArgs a = zeroed<VelhalfArgs>();
        if (!m->vel_same) {
            vh = 3 - m->v0 - m->v1;
            if (vh < 0 || vh > 2 || vh == m->v0 || vh == m->v1) {
                for (vh = 0; vh < 3 && (vh == m->v0 || vh == m->v1); ++vh) {}
            }
            for (int c = 0; c < 3; ++c) { a.a[c] = m->vel[m->v0][c]; a.b[c] = m->vel[m->v1][c]; a.h[c] = m->vel[vh][c]; }
            a.n = m->ncell; a.st = st;
        }
        P.add(K_VELHALF, "velhalf", m->vel_same ? 0 : grid_for(m->ncell, 256, 4096), a);
    }
    const Vel V0{m->vel[m->v0][0], m->vel[m->v0][1], m->vel[m->v0][2]};
    const Vel VH{m->vel[vh][0], m->vel[vh][1], m->vel[vh][2]};
    const Vel V1{m->vel[m->v1][0], m->vel[m->v1][1], m->vel[m->v1][2]};
    /* division marks + capacity check */
    {
        MarksArgs a = zeroed<MarksArgs>();
        a.st = st; a.sflag = m->sflag; a.bmA = m->bm; a.bmD = m->bm + (size_t)m->nsp * m->nw;
        a.nsp = m->nsp; a.nw = m->nw; a.cap = m->cap;
        P.add(K_MARKS, "marks+scan", grid_for((long long)m->nsp * m->nw, 256, 8192), a);
        plan_scan(P, "marks+scan", scan_bits(m->bm + (size_t)m->nsp * m->nw, m->bm, m->nsp, m->nw, m->cap), m->partials, m->G);
        /* abiotic particles_from_bcs!: the cells that create a particle, the free
         * slots, and their capacity check, before anything changes */
        AbioArgs ab = abio_args(m, dt, V0, VH, V1);
        const long long nc = (long long)m->nsa * m->ncell, ns = (long long)m->nsa * m->cap;
        const bool pbc = ab.bcm != 0;
        P.add(K_ABIO_MARK, "abiotic", pbc ? grid_for(nc, 256, 8192) : 0, ab);
        P.add(K_ABIO_FREE, "abiotic", pbc ? grid_for(ns, 256, 8192) : 0, ab);
        plan_scan(P, "abiotic", scan_plain(m->acflag, (uint32_t)(pbc ? nc : 0)), m->apart, m->acpre);
        plan_scan(P, "abiotic", scan_plain(m->afflag, (uint32_t)(pbc ? ns : 0)), m->apart, m->afpre);
        P.add(K_ABIO_LIST, "abiotic", pbc ? grid_for(ns, 256, 8192) : 0, ab);
        P.add(K_ABIO_CHECK, "abiotic", pbc ? 1 : 0, ab);
        CapArgs c = zeroed<CapArgs>();
        c.st = st; c.G = m->G; c.nsp = m->nsp; c.nw = m->nw;
        P.add(K_CAPCHECK, "marks+scan", 1, c);
        /* abiotic particles: creation at the faces, then advection, the random
         * walk, find_inds! and num (time_step.jl:21-38, 152-159) */
        P.add(K_ABIO_SPAWN, "abiotic", pbc ? grid_for(nc, 256, 8192) : 0, ab);
        P.add(K_ABIO_ADV, "abiotic", m->nsa && !g.gm ? grid_for(ns, 256, 8192) : 0, ab);
        P.add(K_ABIO_ADV_GM, "abiotic", m->nsa && g.gm ? grid_for(ns, 256, 8192) : 0, ab);
    }
    const bool land = g.has_land != 0;
    /* Periodic with N < H: the fills read halo planes of the slab they fill */
    const bool alias = (g.tx == PI_PERIODIC && g.Nx < g.Hx) || (g.ty == PI_PERIODIC && g.Ny < g.Hy);
    /* the fused x + y sweeps where the grid allows them (k_xy), else the
     * separate sweeps with their halo fills */
    const bool fxy = xy_supported(g);
    /* The x and y sweeps of the tracers read only the tracers and vel₁: they run
     * on the side stream, beside advection, binning, PAR, physiology and the
     * particle exchange (which neither read their output nor write the
     * tracers), and k_final joins them.  Measured 2.40 -> 2.35 ms per step on
     * the 256x256x64 run (the main kernels fill the GPU: the overlap is mostly
     * at their tails); forking later (before binning, PAR or k_plk) gains no
     * more. */
    auto add_sweeps = [&]() {
        P.mark(K_FORK);
        P.side(true);
        {
            XYArgs xa = zeroed<XYArgs>();
            xa.g = g; xa.st = st; xa.dt = dt; xa.u = V1.u; xa.v = V1.v; xa.out = m->T2[0];
            xa.run = (int)m->shape[PI_S_XY_RUN];
            const long long xyb = fxy ? xy_blocks(g, xa.run) : 0;
            const bool wide = xy_wide(g);
            P.add(K_XY0, "sweep_xy", land || wide ? 0 : xyb, xa);
            P.add(K_XY1, "sweep_xy", land && !wide ? xyb : 0, xa);
            P.add(K_XYW0, "sweep_xy", !land && wide ? xyb : 0, xa);
            P.add(K_XYW1, "sweep_xy", land && wide ? xyb : 0, xa);
            SweepArgs sa = zeroed<SweepArgs>();
            sa.g = g; sa.st = st; sa.dt = dt;
            sa.vel = V1.u; sa.dim = 0;
            sa.out = m->T1[0]; sa.prev = nullptr;
            const long long xb = fxy ? 0 : sweep_x_blocks(g);
            const bool gm = g.gm != 0;
            P.add(K_SWEEPX0, "sweep_x", !gm && !land ? xb : 0, sa);
            P.add(K_SWEEPX1, "sweep_x", !gm && land ? xb : 0, sa);
            P.add(K_SWEEPX2, "sweep_x", gm && !land ? xb : 0, sa);
            P.add(K_SWEEPX3, "sweep_x", gm && land ? xb : 0, sa);
            FillArgs fa = zeroed<FillArgs>();
            fa.g = g; fa.st = st; fa.nf = PI_NTRACERS; fa.dim = 1; fa.lines_full = 0;
            fa.f = m->T1[0];
            P.add(K_HALOCOPY, "fill", alias ? grid_for((long long)g.ncell * PI_NTRACERS, 256, 4096) : 0, fa);
            P.add(K_FILL, "fill", fxy ? 0 : blocks_for((long long)g.Nx * g.Nz * PI_NTRACERS), fa);
            sa.vel = V1.v; sa.dim = 1;
            sa.out = m->T2[0]; sa.prev = m->T1[0];
            const long long yb = fxy ? 0 : grid_for(sweep_y_items(g), 256, 8192);
            P.add(K_SWEEPY0, "sweep_y", !gm && !land ? yb : 0, sa);
            P.add(K_SWEEPY1, "sweep_y", !gm && land ? yb : 0, sa);
            P.add(K_SWEEPY2, "sweep_y", gm && !land ? yb : 0, sa);
            P.add(K_SWEEPY3, "sweep_y", gm && land ? yb : 0, sa);
            fa.dim = 2;
            fa.f = m->T2[0];
            P.add(K_FILL, "fill", fxy ? 0 : blocks_for((long long)g.Nx * g.Ny * PI_NTRACERS), fa);
        }
        P.side(false);
    };

    add_sweeps();
    /* advection + diffusion + cells */
    {
        AdvArgs aa = zeroed<AdvArgs>();
        aa.g = g; aa.st = st; aa.v0 = V0; aa.vh = VH; aa.v1 = V1; aa.dt = dt;
        aa.sx = std::sqrt(m->bgc[PI_B_kappa_hP] * dt);
        aa.sy = std::sqrt(m->bgc[PI_B_kappa_hP] * dt);
        aa.sz = std::sqrt(m->bgc[PI_B_kappa_vP] * dt);
        aa.seed = m->seed;
        aa.newc = m->newc; aa.stayf = m->stayf; aa.small = m->small; aa.qcap = m->qcap; aa.cnts = m->cnts;
        aa.dstamp = m->dstamp; aa.dcell = m->dcell; aa.cap = m->cap; aa.dbd = m->dbd;
        aa.dbd_ptc = m->mode == PI_MODE_IRONENERGY ? 1 : 0;
        /* up to 16384 blocks by default: with 8.4M individuals (two entries per
         * thread) ~5% faster than 4096 blocks; 32768 is no faster, 65536 much slower */
        const long long ab = grid_for(2 * m->ctot, 256, (int)m->shape[PI_S_ADV_BLOCKS]);
        P.add(K_ADV, "adv", g.gm ? 0 : ab, aa);
        P.add(K_ADV_GM, "adv", g.gm ? ab : 0, aa);
    }
    {
        ScanIn sp = scan_plain(m->stayf, (uint32_t)(2 * m->ctot));      /* stayer prefix over the sorted part */
        sp.dn = (const uint32_t*)((const char*)st + offsetof(DevState, nA));   /* device-side length nA */
        plan_scan(P, "bin", sp, m->partials, m->P);
        BinArgs b = zeroed<BinArgs>();
        b.st = st; b.ncell = m->ncell; b.seg = m->seg; b.segp = m->segp; b.P = m->P; b.stayf = m->stayf;
        b.newc = m->newc; b.small = m->small; b.qcap = m->qcap; b.cnts = m->cnts; b.cnt = m->cnt; b.lng = m->lng;
        b.pkey = m->pkey; b.psrc = m->psrc; b.mix = m->mix; b.big = m->big; b.skey = m->skey; b.ssrc = m->ssrc;
        b.Chl = m->Chl; b.pop = m->pop;
        b.pop_sp = m->bgc[PI_B_shared_graz] == 1.0f ? -1 : m->nsp - 1;
        P.add(K_CELLCOUNT, "bin", grid_for(m->ncell, 256, 8192), b);
        plan_scan(P, "bin", scan_plain(m->cnt, (uint32_t)m->ncell), m->partials, m->seg);
        P.add(K_PLACE_STAY, "bin", pb, b);
        P.add(K_PLACE_SMALL, "bin", pb, b);
        P.add(K_MIXSORT, "bin", cb, b);
        P.add(K_BIGCELLS, "bin", 256, b);
        /* one resident round by default (8 blocks per CU at its register use): ~5 us
         * faster than 8192 blocks */
        P.add(K_CELLSUM, "bin", std::min<long long>(grid_for(m->ncell, 256, 8192), m->shape[PI_S_CELLSUM_PER_CU] * m->ncu), b);
        P.add(K_CELLSUM_LONG, "bin", m->long_blocks, b);
    }
    {
        ParArgs a = zeroed<ParArgs>();
        a.g = g; a.st = st; a.Chl = m->Chl; a.PARF_in = m->PARF_in; a.PARF_out = m->PARF_out;
        a.diag_par = m->diagPAR; a.kc = m->bgc[PI_B_kc]; a.kw = m->bgc[PI_B_kw];
        /* one block per 64 columns (up to 64 levels; deeper grids: one lane per column) */
        P.add(K_PAR, "par", g.Nz <= 64 ? grid_for((long long)g.Nx * g.Ny, 64, 8192) : blocks_for((long long)g.Nx * g.Ny), a);
    }
    /* physiology */
    {
        PhysArgs pa = zeroed<PhysArgs>();
        pa.g = g; pa.st = st; pa.tab = m->tab; pa.cap = m->cap; pa.nsp = m->nsp; pa.params = m->d_params;
        pa.seg = m->seg; pa.pkey = m->pkey; pa.psrc = m->psrc; pa.temp = m->temp; pa.pop = m->pop;
        pa.sp_pop = m->bgc[PI_B_shared_graz] == 1.0f ? 0 : 1;   /* time_step.jl:21 vs :80 */
        pa.bmA = m->bm; pa.bmD = m->bm + (size_t)m->nsp * m->nw; pa.G = m->G; pa.nw = m->nw;
        pa.newc = m->newc; pa.sflag = m->sflag;
        pa.con = m->con; pa.diag_num = m->diag; pa.dt = dt; pa.seed = m->seed;
        pa.dstamp = m->dstamp; pa.dcell = m->dcell; pa.rstamp = m->rstamp; pa.rcons = m->rcons; pa.dbd = m->dbd;
        pa.rdidx = m->rdidx;
        const long long phb = grid_for(2 * m->ctot, 256, (int)m->shape[PI_S_PHYS_BLOCKS]);
        P.add(K_PHYS, "phys", m->mode == PI_MODE_QUOTA ? phb : 0, pa);
        P.add(K_PHYSC, "phys", m->mode == PI_MODE_CARBON ? phb : 0, pa);
        P.add(K_PHYSM, "phys", m->mode == PI_MODE_MACROMOLECULAR ? phb : 0, pa);
        P.add(K_PHYSI, "phys", m->mode == PI_MODE_IRONENERGY ? phb : 0, pa);
        IdleArgs ia = zeroed<IdleArgs>();
        ia.st = st; ia.tab = m->tab; ia.params = m->d_params; ia.cap = m->cap; ia.nsp = m->nsp; ia.dt = dt;
        const long long ib = grid_for(m->ctot, 256, 8192);
        P.add(K_IDLEM, "phys", m->mode == PI_MODE_MACROMOLECULAR ? ib : 0, ia);
        P.add(K_IDLEI, "phys", m->mode == PI_MODE_IRONENERGY ? ib : 0, ia);
    }
    {
        PlkArgs ka = zeroed<PlkArgs>();
        ka.g = g; ka.seg = m->seg; ka.pkey = m->pkey; ka.con = m->con; ka.diag = m->diag; ka.st = st;
        ka.rstamp = m->rstamp; ka.rcons = m->rcons; ka.cap = m->cap;
        for (int q = 0; q < PI_NTRACERS; ++q) ka.plk[q] = m->plk[q] ? m->plk[q] : m->plk[PI_DIC];
        ka.lng = m->lng; ka.lg = m->skey; ka.lm = m->ssrc;
        ka.segp = m->segp;                 /* k_plk copies seg -> segp (the binning above has read segp) */
        /* one cell per thread by default: with a loop over cells each iteration's
         * loads wait for the previous iteration's stores (one counter for both) */
        const bool iron = m->mode == PI_MODE_IRONENERGY;
        const long long kb = grid_for(m->ncell, 256, (int)m->shape[PI_S_PLK_BLOCKS]);
        P.add(K_PLK, "plk", iron ? 0 : kb, ka);
        P.add(K_PLK_LONG, "plk", iron ? 0 : m->long_blocks, ka);
        P.add(K_PLKI, "plk", iron ? kb : 0, ka);
        P.add(K_PLK_LONGI, "plk", iron ? m->long_blocks : 0, ka);
    }
    {   /* phytoplankton-particle interaction, release, ptc and num (time_step.jl:120-159) */
        InteractArgs ia = zeroed<InteractArgs>();
        ia.g = g; ia.st = st; ia.seg = m->seg; ia.rstamp = m->rstamp; ia.rdidx = m->rdidx; ia.params = m->d_params;
        for (int sa = 0; sa < m->nsa; ++sa) for (int q = 0; q < PI_NAPARAMS; ++q) ia.ap[sa][q] = m->aparams[sa][q];
        ia.cap = m->cap; ia.nsp = m->nsp; ia.nsa = m->nsa;
        ia.x = m->ax; ia.y = m->ay; ia.z = m->az; ia.xi = m->axi; ia.yi = m->ayi; ia.zi = m->azi; ia.ac = m->aac;
        ia.num = m->anum; ia.diag_ptc = m->diag_ptc;
        ia.cnt = m->icnt; ia.off = m->ioff; ia.fill = m->ifill; ia.list = m->ilist;
        ia.rflag = m->rflag; ia.rrank = m->rrank; ia.fflag = m->ifflag; ia.fpre = m->ifpre; ia.flist = m->iflist;
        ia.mc = m->mc;
        for (int q = 0; q < PI_MAXPAIRS; ++q)
            for (int j = 0; j < 2; ++j) { ia.pint[q][j] = m->pint[q][j]; ia.prel[q][j] = m->prel[q][j]; }
        ia.dt = dt; ia.seed = m->seed;
        /* every model lists the same records (PI_MAXPAIRS of each kind); absent pairs launch nothing */
        const long long nc = m->ncell, cap = m->cap;
        for (int q = 0; q < PI_MAXPAIRS; ++q) {
            const bool on = q < m->npint;
            ia.q = on ? q : 0;
            P.add(K_INT_ZERO, "interact", on ? grid_for(nc, 256, 8192) : 0, ia);
            P.add(K_INT_COUNT, "interact", on ? grid_for(cap, 256, 8192) : 0, ia);
            plan_scan(P, "interact", scan_plain(m->icnt, (uint32_t)(on ? nc : 0)), m->apart, m->ioff);
            P.add(K_INT_LIST, "interact", on ? grid_for(cap, 256, 8192) : 0, ia);
            P.add(K_INT_CELLS, "interact", on ? grid_for(nc, 256, 8192) : 0, ia);
        }
        for (int q = 0; q < PI_MAXPAIRS; ++q) {
            const bool on = q < m->nprel;
            ia.q = on ? q : 0;
            P.add(K_REL_ZERO, "interact", on ? grid_for(cap, 256, 8192) : 0, ia);
            P.add(K_REL_FLAG, "interact", on ? grid_for(m->ctot, 256, 8192) : 0, ia);
            plan_scan(P, "interact", scan_plain(m->rflag, (uint32_t)(on ? cap : 0)), m->apart, m->rrank);
            plan_scan(P, "interact", scan_plain(m->ifflag, (uint32_t)(on ? cap : 0)), m->apart, m->ifpre);
            P.add(K_REL_CHECK, "interact", on ? 1 : 0, ia);
            P.add(K_REL_LIST, "interact", on ? grid_for(cap, 256, 8192) : 0, ia);
            P.add(K_REL_WRITE, "interact", on ? grid_for(m->ctot, 256, 8192) : 0, ia);
        }
        P.add(K_INT_DIAG, "interact", m->diag_ptc ? grid_for(m->ctot, 256, 8192) : 0, ia);
        P.add(K_INT_NUM, "interact", m->npint + m->nprel > 0 ? grid_for((long long)m->nsa * cap, 256, 8192) : 0, ia);
    }
    P.mark(K_JOIN);
    /* tracers: k_final and the halo fill of the new tracers, after the join */
    {
        FillArgs fa = zeroed<FillArgs>();
        fa.g = g; fa.st = st; fa.nf = PI_NTRACERS;
        FinalArgs fn = zeroed<FinalArgs>();
        fn.g = g; fn.st = st; fn.w = V1.w; fn.dt = dt;
        fn.prev = m->T2[0]; fn.plk = m->plk_slab;
        fn.kh = m->bgc[PI_B_kappa_h]; fn.kv = m->bgc[PI_B_kappa_v];
        fn.wsink_inorg = m->bgc[PI_B_w_sink_inorg]; fn.wsink_org = m->bgc[PI_B_w_sink_org];
        fn.kDOC = m->bgc[PI_B_kDOC]; fn.kPOC = m->bgc[PI_B_kPOC]; fn.kDON = m->bgc[PI_B_kDON];
        fn.kPON = m->bgc[PI_B_kPON]; fn.kDOP = m->bgc[PI_B_kDOP]; fn.kPOP = m->bgc[PI_B_kPOP]; fn.Nit = m->bgc[PI_B_Nit];
        fn.lPOC = m->bgc[PI_B_lambda_POC]; fn.lmin = m->bgc[PI_B_lambda_min]; fn.ldust = m->bgc[PI_B_lambda_dust];
        fn.DFeFrac = m->bgc[PI_B_DFeFrac]; fn.kdiss = m->bgc[PI_B_kdiss]; fn.lig = m->bgc[PI_B_ligand];
        fn.lFe = m->bgc[PI_B_lambda_Fe];
        fn.fill_halo = fxy ? 1 : 0;       /* k_final writes the halo images itself */
        fn.run = (int)m->shape[PI_S_FINAL_RUN];
        fn.bc = m->d_bc;
        bool bcs = false;
        for (int q = 0; q < PI_NTRACERS; ++q) {
            for (int f = 0; f < PI_NBC; ++f) fn.bcm[q] |= (uint8_t)((m->h_bc.kind[q][f] != PI_BCK_NONE) << f);
            bcs |= fn.bcm[q] != 0;
        }
        const long long fb = grid_for(final_items(g, fn.run), 256, 8192);
        /* k_final<n>: bit 0 land, bit 1 flux boundary conditions, bit 2 LatLonGrid,
         * bit 3 IronEnergyMode (DFe and PFe_bio exchange) */
        const int fv = (land ? 1 : 0) | (bcs ? 2 : 0) | (g.gm ? 4 : 0) | (m->mode == PI_MODE_IRONENERGY ? 8 : 0);
        static const KId fk[16] = {K_FINAL0, K_FINAL1, K_FINAL2, K_FINAL3, K_FINAL4, K_FINAL5, K_FINAL6, K_FINAL7,
                                   K_FINAL8, K_FINAL9, K_FINAL10, K_FINAL11, K_FINAL12, K_FINAL13, K_FINAL14, K_FINAL15};
        for (int v = 0; v < 16; ++v) P.add(fk[v], "final", v == fv ? fb : 0, fn);
        /* fill_halo_tracer! of the new tracers: x, then y, then z over whole planes */
        fa.f = nullptr;
        fa.lines_full = 1;
        P.add(K_HALOCOPY, "fill", alias ? grid_for((long long)g.ncell * PI_NTRACERS, 256, 4096) : 0, fa);
        fa.dim = 0; P.add(K_FILL, "fill", fxy ? 0 : blocks_for((long long)g.Py * g.Pz * PI_NTRACERS), fa);
        fa.dim = 1; P.add(K_FILL, "fill", fxy ? 0 : blocks_for((long long)g.Px * g.Pz * PI_NTRACERS), fa);
        fa.dim = 2; P.add(K_FILL, "fill", fxy ? 0 : blocks_for((long long)g.Px * g.Py * PI_NTRACERS), fa);
        /* on the grids k_xy takes (halo 2, no aliased periodic direction) k_final writes the halo images */
    }
    {
        AccumArgs a = zeroed<AccumArgs>();
        a.nq = (int)m->diag_ids.size(); a.n = m->ncell; a.st = st;
        for (int j = 0; j < a.nq; ++j) { a.d[j] = m->diag_tr[j]; a.q[j] = m->diag_ids[j]; }
        P.add(K_ACCUM, "diag", a.nq ? grid_for(m->ncell, 256, 4096) : 0, a);
        FinishArgs f = zeroed<FinishArgs>();
        f.st = st; f.seg = m->seg; f.ncell = m->ncell; f.nsp = m->nsp;
        P.add(K_FINISH, "finish", 1, f);
    }
    /* vel₀ <- vel₁ */
    m->v0 = m->v1;
    m->vel_same = true;
    m->host_iter += 1;
    m->last_dt = dt;
}

/* the plan's launches on the model's stream (with per-group timers when on) */
static pi_status launch_plan(pi_model* m, const Plan& P)
{
    const KInfo* K = kinfo();
    const char* grp = nullptr;
    const bool fork = !m->timing;          /* timed runs keep every launch on one stream */
    for (const Rec& r : P.recs) {
        if (r.kid == K_FORK || r.kid == K_JOIN) {
            if (!fork) continue;
            hipStream_t from = r.kid == K_FORK ? m->stream : m->side, to = r.kid == K_FORK ? m->side : m->stream;
            hipEvent_t ev = r.kid == K_FORK ? m->ev_fork : m->ev_join;
            HIPCHECK(hipEventRecord(ev, from));
            HIPCHECK(hipStreamWaitEvent(to, ev, 0));
            continue;
        }
        if (r.nblk == 0) continue;
        if (m->timing && grp != r.grp) {
            if (grp) tend(m);
            tbegin(m, r.grp);
            grp = r.grp;
        }
        void* argp = (void*)(P.args.data() + r.off);
        HIPCHECK(hipLaunchKernel(K[r.kid].one, dim3(r.nblk), dim3(K[r.kid].threads), &argp, 0,
                                 fork && r.side ? m->side : m->stream));
    }
    if (m->timing && grp) tend(m);
    HIPCHECK(hipGetLastError());
    return PI_OK;
}

static pi_status enqueue_step(pi_model* m, float dt)
{
    build_step(m, dt, m->plan);
    return launch_plan(m, m->plan);
}

pi_status pi_step(pi_model* m, float dt)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    pi_status s = rebuild_if_dirty(m);
    if (s != PI_OK) return s;
    if (m->h_state.err) return fail(PI_ERR_CAPACITY, "model is in an error state from an earlier step");
    s = bc_steps_ok(m, 1);
    if (s != PI_OK) return s;
    return enqueue_step(m, dt);
}

/* Graph replay can be switched off for the whole process with the environment
 * variable PI_NO_GRAPHS=1 (every step then launches eagerly; results are the
 * same), e.g. under tools that intercept graph capture. */
static bool graphs_allowed()
{
    static int v = -1;
    if (v < 0) { const char* e = getenv("PI_NO_GRAPHS"); v = (e && e[0] && e[0] != '0') ? 0 : 1; }
    return v == 1;
}

/* n steps.  Steady steps (vel₀ = vel₁, timing off) replay one captured step
 * graph: the launch sequence and its arguments are identical from step to step
 * because everything a step flips or counts lives in the DevState. */
pi_status pi_run(pi_model* m, float dt, int64_t n)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    pi_status s = rebuild_if_dirty(m);
    if (s != PI_OK) return s;
    if (m->h_state.err) return fail(PI_ERR_CAPACITY, "model is in an error state from an earlier step");
    s = bc_steps_ok(m, n);
    if (s != PI_OK) return s;
    for (int64_t i = 0; i < n; ++i) {
        if (m->timing || !m->vel_same || !graphs_allowed()) {
            s = enqueue_step(m, dt);
            if (s != PI_OK) return s;
            continue;
        }
        if (!m->gexec || m->gdt != dt) {
            drop_graph(m);
            hipGraph_t graph = nullptr;
            HIPCHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
            s = enqueue_step(m, dt);          /* records the launches and advances the host counters */
            hipGraph_t g2 = nullptr;
            const hipError_t e = hipStreamEndCapture(m->stream, &g2);
            graph = g2;
            if (s != PI_OK) { if (graph) (void)hipGraphDestroy(graph); return s; }
            if (e != hipSuccess) { if (graph) (void)hipGraphDestroy(graph); return fail(PI_ERR_HIP, hipGetErrorString(e)); }
            const hipError_t e2 = hipGraphInstantiate(&m->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e2 != hipSuccess) { m->gexec = nullptr; return fail(PI_ERR_HIP, hipGetErrorString(e2)); }
            m->gdt = dt;
        } else {
            m->host_iter += 1;                /* what enqueue_step does on the host for a steady step */
            m->last_dt = dt;
        }
        HIPCHECK(hipGraphLaunch(m->gexec, m->stream));
    }
    return PI_OK;
}

pi_status pi_sync(pi_model* m)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    return sync_state(m);
}

/* ------------------------------------------------------------- downloads */
pi_status pi_download_field(pi_model* m, int32_t sp, int32_t field, void* host, int64_t n)
{
    if (!m || !host) return fail(PI_ERR_ARG, "null argument");
    if (sp < 0 || sp >= m->nsp || field < 0 || field >= PI_NFIELDS) return fail(PI_ERR_ARG, "field id");
    if (field >= PI_F_ex0 && field - PI_F_ex0 >= mode_nex(m->mode)) return fail(PI_ERR_ARG, "field id: not a column of this mode");
    if (n != m->cap) return fail(PI_ERR_ARG, "n must equal capacity");
    pi_status s = rebuild_if_dirty(m);
    if (s != PI_OK) return s;
    s = sync_state(m);
    if (s != PI_OK && s != PI_ERR_CAPACITY) return s;
    launch_store_to_table(m->stream, m->d_state, m->tab, m->cap, 2 * m->ctot);
    const size_t es = (field >= PI_F_xi && field <= PI_F_zi) ? 8 : field == PI_F_ac ? 1 : 4;
    void* tmp = (void*)m->skey;   /* scratch, 8*ctot bytes >= 8*cap */
    launch_table_field(m->stream, m->d_state, m->tab, sp, m->cap, field, m->d_params, m->g, tmp, m->mode);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(host, tmp, es * m->cap, hipMemcpyDeviceToHost, m->stream));
    HIPCHECK(hipStreamSynchronize(m->stream));
    return s;
}

pi_status pi_download_tracer(pi_model* m, int32_t q, float* c)
{
    if (!m || !c || q < 0 || q >= PI_NTRACERS) return fail(PI_ERR_ARG, "bad tracer id");
    pi_status s = sync_state(m);
    if (s != PI_OK && s != PI_ERR_CAPACITY) return s;
    HIPCHECK(hipMemcpy(c, m->h_state.tr[m->h_state.tcur][q], m->ncell * 4, hipMemcpyDeviceToHost));
    return s;
}

pi_status pi_download_grid(pi_model* m, int32_t which, float* h, int64_t n)
{
    if (!m || !h) return fail(PI_ERR_ARG, "null argument");
    pi_status s = sync_state(m);
    if (s != PI_OK && s != PI_ERR_CAPACITY) return s;
    const float* src = nullptr;
    long long cnt = m->ncell;
    switch (which) {
        case PI_G_par: case PI_G_par0: src = m->h_state.par[m->h_state.pcur]; break;
        case PI_G_Chl: src = m->Chl; break;
        case PI_G_pop: src = m->pop; break;
        case PI_G_temp: src = m->temp; break;
        case PI_G_PARF: src = m->PARF_out; cnt = (long long)m->g.Nx * m->g.Ny; break;
        case PI_G_u: src = m->vel[m->v1][0]; break;
        case PI_G_v: src = m->vel[m->v1][1]; break;
        case PI_G_w: src = m->vel[m->v1][2]; break;
        case PI_G_diag_PAR: src = m->diagPAR; break;
        default:
            if (which >= PI_G_plk_DIC && which < PI_G_plk_DIC + PI_NTRACERS) {
                src = m->plk[which - PI_G_plk_DIC];
                if (!src) {
                    if (n != cnt) return fail(PI_ERR_ARG, "size mismatch");
                    memset(h, 0, cnt * 4);
                    return s;
                }
            }
    }
    if (!src) return fail(PI_ERR_ARG, "unknown grid field");
    if (n != cnt) return fail(PI_ERR_ARG, "size mismatch");
    HIPCHECK(hipMemcpy(h, src, cnt * 4, hipMemcpyDeviceToHost));
    return s;
}

pi_status pi_download_diag(pi_model* m, int32_t sp, int32_t which, float* h)
{
    if (!m || !h || sp < 0 || sp >= m->nsp || which < 0 || which >= PI_NDIAG) return fail(PI_ERR_ARG, "diag id");
    pi_status s = sync_state(m);
    if (s != PI_OK && s != PI_ERR_CAPACITY) return s;
    if (which == PI_D_ptc) {   /* zero unless an IronEnergyMode model with abiotic particles */
        if (m->diag_ptc) HIPCHECK(hipMemcpy(h, m->diag_ptc + (long long)sp * m->ncell, m->ncell * 4, hipMemcpyDeviceToHost));
        else memset(h, 0, m->ncell * 4);
        return s;
    }
    HIPCHECK(hipMemcpy(h, m->diag + ((long long)sp * 4 + which) * m->ncell, m->ncell * 4, hipMemcpyDeviceToHost));
    return s;
}

pi_status pi_download_diag_tracer(pi_model* m, int32_t q, float* h)
{
    if (!m || !h) return fail(PI_ERR_ARG, "null argument");
    pi_status s = sync_state(m);
    if (s != PI_OK && s != PI_ERR_CAPACITY) return s;
    for (size_t a = 0; a < m->diag_ids.size(); ++a)
        if (m->diag_ids[a] == q) { HIPCHECK(hipMemcpy(h, m->diag_tr[a], m->ncell * 4, hipMemcpyDeviceToHost)); return s; }
    return fail(PI_ERR_ARG, "tracer not diagnosed");
}

pi_status pi_reset_diags(pi_model* m)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    HIPCHECK(hipStreamSynchronize(m->stream));
    HIPCHECK(hipMemset(m->diagPAR, 0, m->ncell * 4));
    HIPCHECK(hipMemset(m->diag, 0, (size_t)m->nsp * 4 * m->ncell * 4));
    for (float* p : m->diag_tr) HIPCHECK(hipMemset(p, 0, m->ncell * 4));
    if (m->nsa) HIPCHECK(hipMemset(m->anum, 0, (size_t)m->nsa * m->ncell * 4));
    if (m->diag_ptc) HIPCHECK(hipMemset(m->diag_ptc, 0, (size_t)m->nsp * m->ncell * 4));
    return PI_OK;
}

pi_status pi_get_iteration(pi_model* m, int64_t* it)
{
    if (!m || !it) return fail(PI_ERR_ARG, "null argument");
    *it = m->host_iter;
    return PI_OK;
}

pi_status pi_individual_steps(pi_model* m, int64_t* n)
{
    if (!m || !n) return fail(PI_ERR_ARG, "null argument");
    pi_status s = sync_state(m);
    if (s != PI_OK && s != PI_ERR_CAPACITY) return s;
    *n = (int64_t)m->h_state.ind_steps;
    return PI_OK;
}

pi_status pi_count_active(pi_model* m, int32_t sp, int64_t* n)
{
    if (!m || !n || sp < 0 || sp >= m->nsp) return fail(PI_ERR_ARG, "bad argument");
    pi_status s = rebuild_if_dirty(m);
    if (s != PI_OK) return s;
    HIPCHECK(hipMemsetAsync(m->d_count, 0, 8, m->stream));
    launch_count(m->stream, m->d_state, sp, m->d_count, 2 * m->ctot);
    unsigned long long c = 0;
    HIPCHECK(hipMemcpyAsync(&c, m->d_count, 8, hipMemcpyDeviceToHost, m->stream));
    HIPCHECK(hipStreamSynchronize(m->stream));
    *n = (int64_t)c;
    return PI_OK;
}

pi_status pi_enable_timing(pi_model* m, int32_t on)
{
    if (!m) return fail(PI_ERR_ARG, "null model");
    HIPCHECK(hipStreamSynchronize(m->stream));
    m->timing = on != 0;
    m->tcount = 0;
    return PI_OK;
}

/* Per launch group recorded since timing was enabled (or last read): total
 * milliseconds and number of launches, aggregated by name; resets the record. */
pi_status pi_kernel_times(pi_model* m, int32_t* n, const char** names, float* ms, int32_t* counts, int32_t cap)
{
    if (!m || !n) return fail(PI_ERR_ARG, "null argument");
    HIPCHECK(hipStreamSynchronize(m->stream));
    int32_t k = 0;
    for (size_t i = 0; i < m->tcount; ++i) {
        float t = 0.0f;
        HIPCHECK(hipEventElapsedTime(&t, m->timers[i].a, m->timers[i].b));
        int32_t j = 0;
        while (j < k && strcmp(names[j], m->timers[i].name) != 0) ++j;
        if (j == k) {
            if (k == cap) continue;
            names[k] = m->timers[i].name; ms[k] = 0.0f; counts[k] = 0; ++k;
        }
        ms[j] += t;
        counts[j] += 1;
    }
    *n = k;
    m->tcount = 0;
    return PI_OK;
}

pi_status pi_probe_copy(int64_t bytes)
{
    if (bytes <= 0 || bytes % 16) return fail(PI_ERR_ARG, "probe size must be a positive multiple of 16");
    void *a = nullptr, *b = nullptr;
    HIPCHECK(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) { (void)hipFree(a); return fail(PI_ERR_NOMEM, "probe allocation"); }
    e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) {
        launch_probe_copy(nullptr, b, a, bytes / 16);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(a); (void)hipFree(b);
    if (e != hipSuccess) return fail(PI_ERR_HIP, hipGetErrorString(e));
    return PI_OK;
}


/* ================================================================ batches
 * Many models stepped together: at every position of the (common) launch
 * sequence one batched kernel runs the blocks of all members, each block
 * finding its model, its index and its block count in a table.  Members keep
 * their own buffers and device state, so results are the same as stepping
 * each alone; a member's device-side error stops only its own work. */
} /* extern "C" */

struct BSlot {
    uint16_t kid = 0;
    uint8_t side = 0;             /* runs on the batch's side stream */
    uint32_t nblk = 0;            /* blocks of the batched launch */
    void* dargs = nullptr;        /* member m's arguments at m * argsz */
    uint4* dbt = nullptr;         /* per block: (member, block, member's blocks, 0) */
    size_t args_cap = 0, bt_cap = 0;
    void* largs = nullptr;        /* the tables the launches read: dargs / dbt, or a place in `scratch` */
    uint4* lbt = nullptr;
};

/* The device array that pi_batch_run_noalloc keeps its launch tables in (no
 * counterpart in the reference).  One copy per GPU, held from the moment the
 * library's code object is loaded on that GPU until the process ends.  Host
 * code never passes the array's name to HIP: its address on a device comes
 * from hipGetSymbolAddress, once per device (scratch_base). */
#define PI_SCRATCH_BYTES (32u << 20)
#define PI_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned char scratch[PI_SCRATCH_BYTES];
static char* g_scratch_base[PI_SCRATCH_DEVICES];
static std::mutex g_scratch_mu;

/* the address of `scratch` on device dev (made current by the caller) */
static pi_status scratch_base(int dev, char** out)
{
    if (dev < 0 || dev >= PI_SCRATCH_DEVICES) return fail(PI_ERR_ARG, "device index out of range for the static scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_base[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(PI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_base[dev] = (char*)p;
    }
    *out = g_scratch_base[dev];
    return PI_OK;
}

struct pi_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;          /* second stream of the step's fork (default priority) */
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<pi_model*> mem;
    std::vector<Plan> plans;      /* per member: the plan the tables were built from */
    std::vector<Plan> next;
    std::vector<BSlot> slots;
    std::vector<size_t> order;    /* members by decreasing capacity per cell */
    hipGraphExec_t gexec = nullptr;
    bool tables = false;
    bool graphs = true;           /* replay captured steps (off when other threads use the device) */
    long long budget = 16384;     /* blocks per batched launch of a block-stride kernel */
};

static void batch_drop_graph(pi_batch* b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
}

static bool same_plan(const Plan& a, const Plan& b)
{
    if (a.recs.size() != b.recs.size() || a.args.size() != b.args.size()) return false;
    for (size_t i = 0; i < a.recs.size(); ++i)
        if (a.recs[i].kid != b.recs[i].kid || a.recs[i].nblk != b.recs[i].nblk || a.recs[i].off != b.recs[i].off)
            return false;
    return std::memcmp(a.args.data(), b.args.data(), a.args.size()) == 0;
}

/* argument and block tables of every position from the members' plans; with
 * sbase set, all of them are packed into [sbase, sbase + scap) (256-byte
 * aligned pieces, one copy) instead of the batch's own allocations */
static pi_status batch_tables(pi_batch* b, char* sbase = nullptr, size_t scap = 0)
{
    const KInfo* K = kinfo();
    const size_t nm = b->mem.size();
    size_t npos = 0;
    for (const Plan& P : b->next) if (!P.recs.empty()) { npos = P.recs.size(); break; }
    for (const Plan& P : b->next)
        if (!P.recs.empty() && P.recs.size() != npos) return fail(PI_ERR_ARG, "batch: members with different step plans");
    HIPCHECK(hipStreamSynchronize(b->stream));        /* the previous tables may still be read */
    if (b->slots.size() < npos) b->slots.resize(npos);
    b->order.resize(nm);
    for (size_t m = 0; m < nm; ++m) b->order[m] = m;
    std::stable_sort(b->order.begin(), b->order.end(), [&](size_t x, size_t y) {
        const pi_model* a = b->mem[x];
        const pi_model* c = b->mem[y];
        return (double)a->ctot / (double)a->nint > (double)c->ctot / (double)c->nint;
    });
    std::vector<uint8_t> ha;
    std::vector<uint4> hb;
    std::vector<uint32_t> nb(nm);
    std::vector<uint8_t> stage;                       /* sbase mode: every table, packed */
    auto put = [&](const void* src, size_t n) -> size_t {
        const size_t off = (stage.size() + 255) & ~size_t(255);
        stage.resize(off + n);
        std::memcpy(stage.data() + off, src, n);
        return off;
    };
    for (size_t p = 0; p < npos; ++p) {
        BSlot& S = b->slots[p];
        int kid = -1;
        for (const Plan& P : b->next) if (!P.recs.empty()) { kid = P.recs[p].kid; break; }
        for (const Plan& P : b->next)
            if (!P.recs.empty() && P.recs[p].kid != kid) return fail(PI_ERR_ARG, "batch: members with different step plans");
        S.kid = (uint16_t)kid;
        S.side = 0;
        if (kid == K_FORK || kid == K_JOIN) { S.nblk = 0; continue; }
        for (const Plan& P : b->next) if (!P.recs.empty()) { S.side = P.recs[p].side; break; }
        const KInfo& ki = K[kid];
        long long total = 0;
        for (size_t m = 0; m < nm; ++m) {
            nb[m] = b->next[m].recs.empty() ? 0u : b->next[m].recs[p].nblk;
            total += nb[m];
        }
        if (ki.stride && total > b->budget) {      /* share the budget in proportion, >= 1 block each */
            long long t2 = 0;
            for (size_t m = 0; m < nm; ++m) {
                if (!nb[m]) continue;
                nb[m] = (uint32_t)std::max(1ll, (long long)nb[m] * b->budget / total);
                t2 += nb[m];
            }
            total = t2;
        }
        S.nblk = (uint32_t)total;
        if (!total) continue;
        ha.assign(nm * ki.argsz, 0);
        hb.clear();
        hb.reserve((size_t)total);
        for (size_t m = 0; m < nm; ++m) {
            if (!nb[m]) continue;
            const Plan& P = b->next[m];
            std::memcpy(ha.data() + m * ki.argsz, P.args.data() + P.recs[p].off, ki.argsz);
        }
        /* Blocks are dispatched roughly in table order.  The per-cell kernels of
         * crowded cells run one ordered chain per cell, so the members with the
         * most individuals per cell go first: their chains set the length of
         * the launch. */
        const bool chains = kid == K_CELLSUM_LONG || kid == K_PLK_LONG || kid == K_BIGCELLS;
        for (size_t r = 0; r < nm; ++r) {
            const size_t m = chains ? b->order[r] : r;
            for (uint32_t k = 0; k < nb[m]; ++k) hb.push_back(make_uint4((uint32_t)m, k, nb[m], 0u));
        }
        if (sbase) {
            const size_t oa = put(ha.data(), ha.size());
            const size_t ob = put(hb.data(), hb.size() * sizeof(uint4));
            S.largs = sbase + oa;
            S.lbt = (uint4*)(sbase + ob);
            continue;
        }
        if (S.args_cap < ha.size()) {
            if (S.dargs) (void)hipFree(S.dargs);
            S.dargs = nullptr;
            HIPCHECK(hipMalloc(&S.dargs, ha.size()));
            S.args_cap = ha.size();
        }
        if (S.bt_cap < hb.size()) {
            if (S.dbt) (void)hipFree(S.dbt);
            S.dbt = nullptr;
            HIPCHECK(hipMalloc((void**)&S.dbt, hb.size() * sizeof(uint4)));
            S.bt_cap = hb.size();
        }
        HIPCHECK(hipMemcpyAsync(S.dargs, ha.data(), ha.size(), hipMemcpyHostToDevice, b->stream));
        HIPCHECK(hipMemcpyAsync(S.dbt, hb.data(), hb.size() * sizeof(uint4), hipMemcpyHostToDevice, b->stream));
        HIPCHECK(hipStreamSynchronize(b->stream));    /* ha / hb are reused */
        S.largs = S.dargs;
        S.lbt = S.dbt;
    }
    if (sbase) {
        if (stage.size() > scap) {
            char buf[200];
            snprintf(buf, sizeof buf, "batch: the launch tables need %zu bytes, more than the %zu bytes of the static scratch",
                     stage.size(), scap);
            return fail(PI_ERR_NOMEM, buf);
        }
        HIPCHECK(hipMemcpyAsync(sbase, stage.data(), stage.size(), hipMemcpyHostToDevice, b->stream));
        HIPCHECK(hipStreamSynchronize(b->stream));
    }
    for (size_t p = npos; p < b->slots.size(); ++p) { b->slots[p].nblk = 0; b->slots[p].kid = 0; }
    b->plans.swap(b->next);
    b->tables = true;
    return PI_OK;
}

static pi_status batch_launch(pi_batch* b)
{
    const KInfo* K = kinfo();
    for (BSlot& S : b->slots) {
        if (S.kid == K_FORK || S.kid == K_JOIN) {
            hipStream_t from = S.kid == K_FORK ? b->stream : b->side, to = S.kid == K_FORK ? b->side : b->stream;
            hipEvent_t ev = S.kid == K_FORK ? b->ev_fork : b->ev_join;
            HIPCHECK(hipEventRecord(ev, from));
            HIPCHECK(hipStreamWaitEvent(to, ev, 0));
            continue;
        }
        if (!S.nblk) continue;
        void* args[2] = {(void*)&S.largs, (void*)&S.lbt};
        HIPCHECK(hipLaunchKernel(K[S.kid].multi, dim3(S.nblk), dim3(K[S.kid].threads), args, 0,
                                 S.side ? b->side : b->stream));
    }
    HIPCHECK(hipGetLastError());
    return PI_OK;
}

extern "C" {

pi_status pi_batch_create(pi_batch** out)
{
    if (!out) return fail(PI_ERR_ARG, "null argument");
    pi_batch* b = new pi_batch();
    hipError_t e = hipGetDevice(&b->dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (b->stream) (void)hipStreamDestroy(b->stream);
        if (b->side) (void)hipStreamDestroy(b->side);
        if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
        if (b->ev_join) (void)hipEventDestroy(b->ev_join);
        delete b;
        return fail(PI_ERR_HIP, hipGetErrorString(e));
    }
    *out = b;
    return PI_OK;
}

pi_status pi_batch_add(pi_batch* b, pi_model* m)
{
    if (!b || !m) return fail(PI_ERR_ARG, "null argument");
    if (m->batch) return fail(PI_ERR_ARG, "model is already in a batch");
    if (m->dev != b->dev) return fail(PI_ERR_ARG, "model and batch are on different devices");
    HIPCHECK(hipStreamSynchronize(m->stream));
    drop_graph(m);
    m->own_stream = m->stream;
    m->stream = b->stream;
    m->batch = b;
    b->mem.push_back(m);
    b->plans.emplace_back();
    b->tables = false;
    batch_drop_graph(b);
    return PI_OK;
}

pi_status pi_batch_remove(pi_batch* b, pi_model* m)
{
    if (!b || !m) return fail(PI_ERR_ARG, "null argument");
    auto it = std::find(b->mem.begin(), b->mem.end(), m);
    if (it == b->mem.end()) return fail(PI_ERR_ARG, "model is not in this batch");
    HIPCHECK(hipStreamSynchronize(b->stream));
    const size_t k = (size_t)(it - b->mem.begin());
    b->mem.erase(it);
    b->plans.erase(b->plans.begin() + (long)k);
    drop_graph(m);
    m->stream = m->own_stream;
    m->own_stream = nullptr;
    m->batch = nullptr;
    b->tables = false;
    batch_drop_graph(b);
    return PI_OK;
}

/* n steps of every member: each member's next step is planned (staging changes
 * made since the last step included), the tables are rebuilt when a plan
 * changed, and a steady step replays one captured graph of the batched
 * launches. */
pi_status pi_batch_run(pi_batch* b, float dt, int64_t n)
{
    if (!b) return fail(PI_ERR_ARG, "null batch");
    const size_t nm = b->mem.size();
    if (nm == 0) return PI_OK;
    for (pi_model* md : b->mem) {
        pi_status s = bc_steps_ok(md, n);
        if (s != PI_OK) return s;
    }
    b->next.resize(nm);
    for (int64_t i = 0; i < n; ++i) {
        bool same = b->tables;
        for (size_t m = 0; m < nm; ++m) {
            pi_model* md = b->mem[m];
            pi_status s = rebuild_if_dirty(md);
            if (s != PI_OK) return s;
            if (md->h_state.err) { b->next[m].clear(); }
            else build_step(md, dt, b->next[m]);
            if (same && !same_plan(b->next[m], b->plans[m])) same = false;
        }
        if (!same) {
            batch_drop_graph(b);
            pi_status s = batch_tables(b);
            if (s != PI_OK) return s;
        }
        if (!same || !b->graphs || !graphs_allowed()) {   /* a changed step runs without a graph */
            pi_status s = batch_launch(b);
            if (s != PI_OK) return s;
            continue;
        }
        if (!b->gexec) {
            hipGraph_t graph = nullptr;
            HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            pi_status s = batch_launch(b);
            const hipError_t e = hipStreamEndCapture(b->stream, &graph);
            if (s != PI_OK) { if (graph) (void)hipGraphDestroy(graph); return s; }
            if (e != hipSuccess) { if (graph) (void)hipGraphDestroy(graph); return fail(PI_ERR_HIP, hipGetErrorString(e)); }
            const hipError_t e2 = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e2 != hipSuccess) { b->gexec = nullptr; return fail(PI_ERR_HIP, hipGetErrorString(e2)); }
        }
        HIPCHECK(hipGraphLaunch(b->gexec, b->stream));
    }
    return PI_OK;
}

/* pi_batch_run with its launch tables in the static device array `scratch`
 * instead of device allocations: no hipMalloc / hipFree / graph instantiation
 * (every step is launched eagerly).  The tables are written at every call
 * and whenever a member's plan changes; calls of this function on one GPU
 * share that GPU's array, so calls that overlap in time on one GPU race
 * (serialise them).  Tables larger than PI_SCRATCH_BYTES are refused with
 * PI_ERR_NOMEM before anything is launched. */
pi_status pi_batch_run_noalloc(pi_batch* b, float dt, int64_t n)
{
    if (!b) return fail(PI_ERR_ARG, "null batch");
    const size_t nm = b->mem.size();
    if (nm == 0 || n <= 0) return PI_OK;
    for (pi_model* md : b->mem) {
        pi_status s = bc_steps_ok(md, n);
        if (s != PI_OK) return s;
    }
    int cur = 0;
    HIPCHECK(hipGetDevice(&cur));
    if (cur != b->dev) HIPCHECK(hipSetDevice(b->dev));
    char* base = nullptr;
    pi_status st = scratch_base(b->dev, &base);
    if (cur != b->dev) (void)hipSetDevice(cur);
    if (st != PI_OK) return st;
    batch_drop_graph(b);
    b->next.resize(nm);
    bool fresh = true;                /* this call has not written the scratch tables yet */
    for (int64_t i = 0; i < n; ++i) {
        bool same = !fresh;
        for (size_t m = 0; m < nm; ++m) {
            pi_model* md = b->mem[m];
            pi_status s = rebuild_if_dirty(md);
            if (s != PI_OK) return s;
            if (md->h_state.err) { b->next[m].clear(); }
            else build_step(md, dt, b->next[m]);
            if (same && !same_plan(b->next[m], b->plans[m])) same = false;
        }
        if (!same) {
            pi_status s = batch_tables(b, base, PI_SCRATCH_BYTES);
            if (s != PI_OK) { b->tables = false; return s; }
            fresh = false;
        }
        pi_status s = batch_launch(b);
        if (s != PI_OK) { b->tables = false; return s; }
    }
    b->tables = false;                /* pi_batch_run rebuilds its own tables */
    return PI_OK;
}

pi_status pi_batch_set_budget(pi_batch* b, int64_t blocks)
{
    if (!b) return fail(PI_ERR_ARG, "null batch");
    if (blocks < 256 || blocks > (1ll << 22)) return fail(PI_ERR_ARG, "budget out of range [256, 2^22]");
    if (b->budget != blocks) {
        b->budget = blocks;
        b->tables = false;             /* rebuilt at the next step */
        batch_drop_graph(b);
    }
    return PI_OK;
}

pi_status pi_batch_graphs(pi_batch* b, int32_t on)
{
    if (!b) return fail(PI_ERR_ARG, "null batch");
    b->graphs = on != 0;
    if (!b->graphs) batch_drop_graph(b);
    return PI_OK;
}

pi_status pi_batch_sync(pi_batch* b)
{
    if (!b) return fail(PI_ERR_ARG, "null batch");
    HIPCHECK(hipStreamSynchronize(b->stream));
    return PI_OK;
}

pi_status pi_batch_size(const pi_batch* b, int32_t* n)
{
    if (!b || !n) return fail(PI_ERR_ARG, "null argument");
    *n = (int32_t)b->mem.size();
    return PI_OK;
}

pi_status pi_batch_destroy(pi_batch* b)
{
    if (!b) return PI_OK;
    (void)hipStreamSynchronize(b->stream);
    while (!b->mem.empty()) (void)pi_batch_remove(b, b->mem.back());
    batch_drop_graph(b);
    for (BSlot& S : b->slots) { if (S.dargs) (void)hipFree(S.dargs); if (S.dbt) (void)hipFree(S.dbt); }
    (void)hipStreamDestroy(b->stream);
    (void)hipStreamDestroy(b->side);
    (void)hipEventDestroy(b->ev_fork);
    (void)hipEventDestroy(b->ev_join);
    delete b;
    return PI_OK;
}

}  /* extern "C" */
