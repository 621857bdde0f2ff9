/*
 * k_args.h - kernel argument blocks and the host-side launchers of each kernel
 * file (launchers live next to their kernels).
 */
#pragma once
#include "pi_device.h"

struct AdvArgs {
    DevGrid g;
    DevState* st;
    Vel v0, vh, v1;
    float dt, sx, sy, sz;
    uint64_t seed;
    uint32_t* newc;           /* per store entry: its cell in this step */
    uint32_t* stayf;          /* per sorted entry: 1 = keeps its cell */
    uint32_t* small;          /* queue of the other live entries: PI_NQ shards of qcap entries */
    uint32_t qcap;
    uint32_t* cnts;           /* per cell: queued arrivals */
    int32_t* dstamp;          /* per slot: iteration in which its occupant dies */
    uint32_t* dcell;          /* per slot: that occupant's cell in this step */
    float* dbd;               /* CarbonMode, per slot: that occupant's Bd (null in QuotaMode) */
    long long cap;
    int dbd_ptc;              /* IronEnergyMode: stash the occupant's ptc (extra column 7) instead */
};

struct PhysArgs {
    DevGrid g;
    DevState* st;
    SlotTable tab;            /* per species rows at tab + sp*cap */
    long long cap;
    int nsp;
    const float* params;      /* nsp * PI_NPARAMS */
    const uint32_t* seg;      /* seg[ncell] = number of live individuals */
    const uint32_t* pkey;
    const uint32_t* psrc;
    const float *temp, *pop;
    int sp_pop;               /* 1: pop per species (shared_graz != 1), counted in the cell's segment */
    const uint32_t* bmA;
    const uint32_t* bmD;
    const uint32_t* G;        /* exclusive scan over [D words | free words] of all species */
    int nw;
    const uint32_t* newc;     /* per store entry: its cell in this step */
    uint8_t* sflag;           /* per slot: flags for the next step's division marks */
    Contrib con;
    float* diag_num;          /* diags: species stride 4*ncell, field num */
    float dt;
    uint64_t seed;
    const int32_t* dstamp;
    const uint32_t* dcell;
    int32_t* rstamp;          /* per slot: iteration in which a daughter was born there */
    float* rcons;             /* per slot x 5: that daughter's consumption */
    const float* dbd;         /* CarbonMode, per slot: Bd of an occupant dying in this step
                               * (IronEnergyMode: its ptc) */
    uint32_t* rdidx;          /* per slot: the next-store index of a daughter reborn there
                               * (models with phytoplankton-particle interaction; else null) */
};

struct PlkArgs {
    DevGrid g;
    const uint32_t* seg;
    const uint32_t* pkey;
    Contrib con;
    float* plk[PI_NTRACERS];
    float* diag;              /* [sp][4][ncell] */
    DevState* st;
    const int32_t* rstamp;
    const float* rcons;
    long long cap;
    const uint32_t* lng;      /* the long cells of this step (k_plk_long) */
    uint32_t* segp;           /* receives this step's segments (the next step's previous ones) */
    uint32_t *lg, *lm;        /* scratch: ordered lists of grazed / dead entries */
};

/* Tracer sets are slabs of PI_NTRACERS padded fields: field q at slab + q·ncell. */
struct SweepArgs {
    DevGrid g;
    DevState* st;
    const float* vel;                 /* u (x sweep) or v (y sweep) of vel₁ */
    float* out;                       /* slab: temp after this sweep */
    const float* prev;                /* slab: temp before this sweep (y sweep) */
    float dt;
    int dim;                          /* 0 = x, 1 = y */
};

/* fused x + y sweeps: reads the current tracers, writes temp after the y sweep */
struct XYArgs {
    DevGrid g;
    DevState* st;
    const float* u;                   /* vel₁ u */
    const float* v;                   /* vel₁ v */
    float* out;                       /* slab: temp after the y sweep (z halos of interior lines filled) */
    float dt;
    int run;                          /* rows of one level per work item (launch shape, any value >= 1) */
};

struct FillArgs {
    DevGrid g;
    DevState* st;
    float* f;                  /* slab; null = the tracers written by this step */
    int dim, lines_full, nf;
};

struct FinalArgs {
    DevGrid g;
    DevState* st;
    const float* w;                   /* vel₁ w */
    const float* prev;                /* slab: temp after the y sweep (z halos filled) */
    const float* plk;                 /* slab of plk fields (read for the tracers with particle exchange) */
    float dt;
    float kh, kv;                     /* κh, κv (tracer_diffusion!) */
    float wsink_inorg, wsink_org;
    float kDOC, kPOC, kDON, kPON, kDOP, kPOP, Nit;
    float lPOC, lmin, ldust, DFeFrac, kdiss, lig, lFe;
    int fill_halo;                    /* also write the halo cells of the new tracers (halo 2, no aliasing) */
    int run;                          /* levels per column run (launch shape, any value >= 1) */
    const struct BcTab* bc;           /* flux boundary conditions (the k_final2/3 instantiations) */
    uint8_t bcm[PI_NTRACERS];         /* per tracer, bit f set = a boundary condition on face f */
};

/* Flux boundary conditions of the tracers (set_bc!, boundary_conditions.jl:12-29),
 * in device memory: per tracer and face (PI_BC_WEST..PI_BC_TOP) getbc's kind,
 * the Number, and the face array(s) (column-major, face size x nt) */
struct BcTab {
    int32_t kind[PI_NTRACERS][PI_NBC];
    float val[PI_NTRACERS][PI_NBC];
    const float* arr[PI_NTRACERS][PI_NBC];
};

/* Abiotic particles (Individuals/Abiotic): per species sa, `cap` slots at
 * sa·cap.  The particle flux boundary conditions (set_bc_particle!) per species
 * and face, in device memory like BcTab. */
struct AbioBc {
    int32_t kind[PI_MAXSA][PI_NBC];
    float val[PI_MAXSA][PI_NBC];
    const float* arr[PI_MAXSA][PI_NBC];
};
/* error species code of a failed abiotic capacity check: PI_ERR_SP_ABIO + sa */
#define PI_ERR_SP_ABIO 1000u
#define PI_ERR_SP_RELEASE 2000u   /* + abiotic species: particle_release! found too few free slots */
struct AbioArgs {
    AdvArgs adv;                      /* grid, state, velocities, Δt, walk scales, stream seed */
    int nsa;
    long long cap, ncell;
    float *x, *y, *z;                 /* [nsa][cap] */
    int32_t *xi, *yi, *zi;            /* find_inds! of the previous step (0 before the first) */
    uint8_t* ac;
    uint32_t* cflag;                  /* [nsa][ncell]: 1 = a particle is created at this cell */
    const uint32_t* cpre;             /* its exclusive prefix, total at [nsa·ncell] */
    uint32_t* fflag;                  /* [nsa][cap]: 1 = free slot */
    const uint32_t* fpre;
    uint32_t* flist;                  /* [nsa][cap]: the free slots of each species in ascending order */
    float* num;                       /* [nsa][ncell]: diags.abiotics.saN.num */
    int num_late;                     /* 1: num is counted after the interactions (k_int_num) */
    const AbioBc* bc;
    uint32_t bcm;                     /* bit 6·sa + f: species sa has a condition on face f */
    float p[PI_MAXSA][PI_NAPARAMS];
};

struct ScanIn {
    const uint32_t* a;
    uint32_t n;               /* length (upper bound when dn is set) */
    const uint32_t* dn;       /* optional device-side length <= n; out[*dn] = total */
    int bits;
    const uint32_t* D;
    const uint32_t* A;
    int nsp, nw;
    long long cap;
};

struct MarksArgs {
    const DevState* st;
    const uint8_t* sflag;
    uint32_t *bmA, *bmD;
    int nsp, nw;
    long long cap;
};

/* binning: cell counts, placement, re-sorting, per-cell Chl and pop */
struct BinArgs {
    DevState* st;
    long long ncell;
    const uint32_t* seg;      /* this step's segments */
    const uint32_t* segp;     /* the previous step's segments */
    const uint32_t* P;        /* stayer prefix over the previous order */
    const uint32_t* stayf;
    const uint32_t* newc;
    const uint32_t* small;
    uint32_t qcap;
    uint32_t* cnts;
    uint32_t* cnt;
    uint32_t* lng;
    uint32_t *pkey, *psrc;
    uint32_t* mix;
    uint32_t* big;
    uint32_t *skey, *ssrc;
    float *Chl, *pop;
    int pop_sp;               /* -1: pop counts every species (shared_graz = 1); else only this species */
};

struct ParArgs {
    DevGrid g;
    DevState* st;
    const float* Chl;
    const float* PARF_in;
    float *PARF_out, *diag_par;
    float kc, kw;
};

struct ScanArgs {
    ScanIn in;
    uint32_t* partials;
    uint32_t* out;
    uint32_t nb;              /* tiles */
};

struct CapArgs { DevState* st; const uint32_t* G; int nsp, nw; };
struct FinishArgs { DevState* st; const uint32_t* seg; long long ncell; int nsp; };
struct VelhalfArgs { const float* a[3]; const float* b[3]; float* h[3]; long long n; const DevState* st; };
struct AccumArgs { float* d[PI_NTRACERS]; int q[PI_NTRACERS]; int nq; long long n; const DevState* st; };

/* The kernels of one step.  Each has a single-model entry point (arguments by
 * value) and a batched one (a table of arguments, one per model, and a table
 * of (model, block, block count) per block). */
enum KId {
    K_VELHALF, K_MARKS, K_SCAN_P, K_SCAN_F, K_CAPCHECK, K_ADV, K_CELLCOUNT, K_PLACE_STAY,
    K_PLACE_SMALL, K_MIXSORT, K_BIGCELLS, K_CELLSUM, K_CELLSUM_LONG, K_PAR, K_PHYS, K_PLK, K_PLK_LONG,
    K_SWEEPX0, K_SWEEPX1, K_XY0, K_XY1, K_SWEEPY0, K_SWEEPY1, K_HALOCOPY, K_FILL, K_FINAL0, K_FINAL1, K_ACCUM,
    K_FINISH, K_XYW0, K_XYW1, K_PHYSC, K_FINAL2, K_FINAL3,
    K_ADV_GM, K_SWEEPX2, K_SWEEPX3, K_SWEEPY2, K_SWEEPY3, K_FINAL4, K_FINAL5, K_FINAL6, K_FINAL7,
    K_ABIO_MARK, K_ABIO_FREE, K_ABIO_LIST, K_ABIO_CHECK, K_ABIO_SPAWN, K_ABIO_ADV, K_ABIO_ADV_GM,
    K_PHYSM, K_IDLEM, K_PHYSI, K_IDLEI, K_PLKI, K_PLK_LONGI,
    K_INT_ZERO, K_INT_COUNT, K_INT_LIST, K_INT_CELLS, K_REL_ZERO, K_REL_FLAG, K_REL_CHECK, K_REL_LIST,
    K_REL_WRITE, K_INT_DIAG, K_INT_NUM, K_FINAL8, K_FINAL9, K_FINAL10, K_FINAL11, K_FINAL12, K_FINAL13,
    K_FINAL14, K_FINAL15, K_NKID
};
/* plan markers (not kernels): fork to / join from the side stream */
enum { K_FORK = K_NKID, K_JOIN = K_NKID + 1 };
struct KInfo {
    const void* one;          /* __global__ void NAME(ARGS) */
    const void* multi;        /* __global__ void NAME_m(const ARGS*, const uint4*) */
    uint32_t argsz;
    uint32_t threads;
    bool stride;              /* the body loops over its work with a block stride (any block count works) */
    const char* name;
};
#define PI_REG(K, ID, NAME, ARGS, THREADS, STRIDE) \
    K[ID] = KInfo{(const void*)&NAME, (const void*)&NAME##_m, (uint32_t)sizeof(ARGS), THREADS, STRIDE, #NAME}
void reg_particles(KInfo* k);
void reg_tracers(KInfo* k);
void reg_util(KInfo* k);

/* grid sizes of the kernels whose work is one item per thread */
long long sweep_x_blocks(const DevGrid& g);
long long sweep_y_items(const DevGrid& g);
long long final_items(const DevGrid& g, int run);
bool xy_supported(const DevGrid& g);   /* k_xy handles this grid */
bool xy_wide(const DevGrid& g);        /* with the wide blocks (k_xyw) */
long long xy_blocks(const DevGrid& g, int run);
uint32_t scan_tiles(uint32_t n);

/* initial state generated on the device (initialize_plankton!,
 * QuotaMode/plankton_generation.jl:35-72, and generate_tracers,
 * tracer_fields.jl:93-134, with the shared-stream draws) */
struct InitArgs {
    DevGrid g;
    const DevState* st;
    SlotTable tab;
    const float* params;
    long long cap;
    int nsp;
    long long N[PI_MAXSP];
    uint64_t seed;
    int mode;                 /* PI_MODE_QUOTA / PI_MODE_CARBON / PI_MODE_MACROMOLECULAR */
};
/* phytoplankton-particle interaction and release (Abiotic/particle_interaction.jl:1-157,
 * time_step.jl:120-159): one workgroup per model runs the pairs in order */
#define PI_MAXPAIRS 4     /* palat.intac and palat.release pairs per model */
#define PI_MAXCAND 64     /* max_candidates */
struct InteractArgs {
    DevGrid g;
    DevState* st;
    const uint32_t* seg;              /* this step's cell segments of the sorted input */
    const int32_t* rstamp;            /* per slot: iteration in which a daughter was reborn there */
    const uint32_t* rdidx;            /* per slot: that daughter's index in the next store */
    const float* params;              /* nsp * PI_NPARAMS (max_ptc) */
    float ap[PI_MAXSA][PI_NAPARAMS];
    long long cap;
    int nsp, nsa;
    float *x, *y, *z;                 /* abiotic [nsa][cap] */
    int32_t *xi, *yi, *zi;
    uint8_t* ac;
    float* num;                       /* [nsa][ncell] */
    float* diag_ptc;                  /* [nsp][ncell]: diags.phytos.spN.ptc (IronEnergyMode; else null) */
    uint32_t *cnt, *off, *fill, *list;            /* per cell counts / offsets (+ total), the cells' particle lists */
    uint32_t *rflag, *rrank, *fflag, *fpre, *flist;   /* per slot: releases, free abiotic slots (ranks + totals) */
    int q;                            /* the pair this launch handles */
    int mc;
    int pint[PI_MAXPAIRS][2], prel[PI_MAXPAIRS][2];   /* (species, abiotic species) */
    float dt;
    uint64_t seed;
};

/* MacroMolecularMode: the growth of the slots that are inactive through a
 * whole step, whose state keeps changing (biosynthesis is not masked by ac) */
struct IdleArgs {
    const DevState* st;
    SlotTable tab;
    const float* params;
    long long cap;
    int nsp;
    float dt;
};
struct TracerInitArgs {
    DevGrid g;
    float* slab;
    float c0[PI_NTRACERS];
    float lower[PI_NTRACERS];
    int32_t n[PI_NTRACERS];
    float step;
    uint64_t seed;
};
void launch_init_individuals(hipStream_t s, const InitArgs& a);
void launch_init_tracers(hipStream_t s, const TracerInitArgs& a);

/* k_util.hip: kernels outside the step (uploads, downloads, measurement) */
void launch_zero(hipStream_t s, uint32_t* p, long long n, const DevState* st);
void launch_store_to_table(hipStream_t s, DevState* st, const SlotTable& tab, long long cap, long long maxn);
void launch_rebuild(hipStream_t s, DevState* st, const SlotTable& tab, long long cap, int nsp, uint8_t* sflag);
void launch_probe_copy(hipStream_t s, void* dst, const void* src, long long n16);
void launch_table_field(hipStream_t s, const DevState* st, const SlotTable& tab, int sp, long long cap, int field,
                        const float* params, const DevGrid& g, void* out, int mode);
void launch_count(hipStream_t s, const DevState* st, int sp, unsigned long long* out, long long maxn);
ScanIn scan_plain(const uint32_t* a, uint32_t n);
ScanIn scan_bits(const uint32_t* D, const uint32_t* A, int nsp, int nw, long long cap);
