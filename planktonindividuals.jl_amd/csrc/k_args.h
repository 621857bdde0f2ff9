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
    uint32_t* small;          /* queue of the other live entries */
    uint32_t* cnts;           /* per cell: queued arrivals */
    int32_t* dstamp;          /* per slot: iteration in which its occupant dies */
    uint32_t* dcell;          /* per slot: that occupant's cell in this step */
    long long cap;
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

/* k_particles.hip */
void launch_adv(hipStream_t st, int blocks, const AdvArgs& a);
void launch_marks(hipStream_t s, const DevState* st, const uint8_t* sflag, uint32_t* bmA, uint32_t* bmD,
                  int nsp, int nw, long long cap);
void launch_cellcount(hipStream_t s, DevState* st, long long ncell, const uint32_t* segp, const uint32_t* P,
                      const uint32_t* cnts, uint32_t* cnt, uint32_t* lng);
void launch_place_stay(hipStream_t s, int blocks, const DevState* st, const uint32_t* seg, const uint32_t* segp,
                       const uint32_t* P, const uint32_t* stayf, const uint32_t* newc, uint32_t* pkey, uint32_t* psrc);
void launch_place_small(hipStream_t s, int blocks, DevState* st, const uint32_t* seg, const uint32_t* small,
                        const uint32_t* newc, uint32_t* cnts, uint32_t* pkey, uint32_t* psrc, uint32_t* mix);
void launch_mixsort(hipStream_t s, int blocks, DevState* st, const uint32_t* seg, const uint32_t* mix,
                    uint32_t* pkey, uint32_t* psrc, uint32_t* big);
void launch_bigcells(hipStream_t s, int blocks, DevState* st, const uint32_t* big, const uint32_t* seg,
                     const uint32_t* segp, const uint32_t* P, uint32_t* pkey, uint32_t* psrc, uint32_t* skey,
                     uint32_t* ssrc);
void launch_cellsum(hipStream_t s, const DevState* st, long long ncell, const uint32_t* seg, const uint32_t* psrc,
                    float* Chl, float* pop, const uint32_t* lng, int long_blocks);
void launch_par(hipStream_t s, int blocks, DevState* st, const DevGrid& g, const float* Chl, const float* PARF_in,
                float* PARF_out, float* diag_par, float kc, float kw);
void launch_phys(hipStream_t st, int blocks, const PhysArgs& a);
void launch_plk(hipStream_t st, int blocks, const PlkArgs& a, const uint32_t* lng, uint32_t* lg, uint32_t* lm,
                int long_blocks);
/* k_tracers.hip */
void launch_sweep(hipStream_t s, int blocks, const SweepArgs& a);
void launch_fill(hipStream_t s, int blocks, const FillArgs& a);
void launch_halo_copy(hipStream_t s, const FillArgs& a);
void launch_final(hipStream_t s, int blocks, const FinalArgs& a);
/* k_util.hip */
void launch_scan(hipStream_t s, const ScanIn& in, uint32_t* partials, uint32_t* out);
ScanIn scan_plain(const uint32_t* a, uint32_t n);
ScanIn scan_bits(const uint32_t* D, const uint32_t* A, int nsp, int nw, long long cap);
void launch_capcheck(hipStream_t s, DevState* st, const uint32_t* G, int nsp, int nw);
void launch_finish(hipStream_t s, DevState* st, const uint32_t* seg, long long ncell, int nsp);
void launch_zero(hipStream_t s, uint32_t* p, long long n, const DevState* st);
void launch_velhalf(hipStream_t s, const float* a, const float* b, float* h, long long n, const DevState* st);
void launch_accum(hipStream_t s, float* d, const DevState* st, int q, long long n);
void launch_store_to_table(hipStream_t s, DevState* st, const SlotTable& tab, long long cap, long long maxn);
void launch_rebuild(hipStream_t s, DevState* st, const SlotTable& tab, long long cap, int nsp, uint8_t* sflag);
void launch_copy_u32(hipStream_t s, uint32_t* dst, const uint32_t* src, long long n, const DevState* st);
void launch_probe_copy(hipStream_t s, void* dst, const void* src, long long n16);
void launch_table_field(hipStream_t s, const DevState* st, const SlotTable& tab, int sp, long long cap, int field,
                        const float* params, const DevGrid& g, void* out);
void launch_count(hipStream_t s, const DevState* st, int sp, unsigned long long* out, long long maxn);
