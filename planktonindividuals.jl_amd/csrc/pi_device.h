/*
 * pi_device.h - device-side data layout and shared helpers of libpi_hip.so.
 *
 * HBM layout (one model):
 *   - grid fields: padded (Nx+2Hx, Ny+2Hy, Nz+2Hz) float32, x fastest;
 *     14 tracers double-buffered, sweep scratch 2 x 14, vel0/vel1/velhalf,
 *     par/par0/Chl/pop/temp/plk/diagnostics.
 *   - particle STORE: only live individuals, all species, SoA, kept sorted by
 *     (cell, species, slot) from one step to the next.  Key = species << 27 |
 *     slot.  Neighbouring lanes therefore touch neighbouring cells: velocity and
 *     tracer gathers hit L1/L2, and every per-cell sum is a contiguous segment.
 *   - SLOT TABLE: the reference's by-slot layout (per species, length capacity),
 *     written on upload and when an individual dies; downloads merge it with
 *     the store.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/pi_hip.h"
#include "../../include/pi_math.h"

#define PI_SLOT_BITS 27
#define PI_SLOT_MASK ((1u << PI_SLOT_BITS) - 1u)

/* flags byte of a store entry / slot-table row */
#define F_AC    1u   /* active */
#define F_GRAZ  2u   /* grazing decision of the previous step */
#define F_MORT  4u   /* mortality decision */
#define F_DVID  8u   /* division decision */
#define F_ZERO  16u  /* stale (xi,yi,zi) = (0,0,0): initial state of the reference */

/* contrib tags */
#define T_ALIVE 1u
#define T_GRAZ  2u
#define T_MORT  4u
#define T_DIV   8u

struct DevGrid {
    int Nx, Ny, Nz, Hx, Hy, Hz, Px, Py, Pz;
    int tx, ty, tz;
    int has_land;
    float dx, dy;
    long long ncell;
    const float* dzC;   /* Pz, 0-based (reference index k -> dzC[k-1]) */
    const float* dzF;
    const uint8_t* lm;  /* padded landmask */
    /* reciprocals of the fixed divisors, RN_double(1/b), for divc */
    double ydx, ydy, y6;
    const double* ydzC;  /* Pz */
    const double* ydzF;  /* Pz */
    const double* yvol;  /* Pz */
    const float* vol;    /* Pz: (Δx·Δy)·dzF[k], the cell volume as the reference forms it */
    /* LatLonGrid (gm = 1): the spacings, areas and volumes per cell
     * (Grids/lat_lon_grid.jl:1-37), padded (Px, Py) and (Px, Py, Pz) arrays;
     * every quotient by them is an IEEE division */
    int gm;
    const float *mdxC, *mdyC, *mdxF, *mdyF, *mAz, *mAx, *mAy, *mVol;
};

/* a / b correctly rounded, without branches, given yd = RN_double(1/b).
 * In binary64, q0 = RN(a·yd) is within an ulp of a/b and yd within half an
 * ulp of 1/b, so one Markstein step (exact residual r = a − b·q0 by FMA, then
 * RN(q0 + r·yd)) gives RN_double(a/b); no step can underflow or overflow in
 * binary64 for float operands.  Rounding RN_double(a/b) to float is RN_float(a/b):
 * 53 >= 2·24 + 2 (double rounding is innocuous for division), and an exact
 * quotient that is a float rounding boundary is representable in binary64.
 * When r is zero or NaN (exact quotient, a or b zero or infinite) q0 is already
 * the IEEE result, signed zeros and infinities included. */
__device__ __forceinline__ float divc(float a, float b, double yd)
{
    const double ad = (double)a, bd = (double)b;
    const double q0 = ad * yd;
    const double r = __builtin_fma(-bd, q0, ad);
    const double q = __builtin_fma(r, yd, q0);
    return (float)((r < 0.0 || r > 0.0) ? q : q0);
}

struct Store {
    uint32_t* key;
    float *x, *y, *z, *Sz, *Bm, *Cq, *Nq, *Pq, *Chl, *gen, *age;
    uint8_t* flags;
    uint32_t* cell;
    float* ex[PI_NEX];           /* extra state columns of the modes that have them (else null) */
};
/* extra state columns per individual of a physiology mode */
__host__ __device__ constexpr int mode_nex(int mode)
{
    return mode == PI_MODE_MACROMOLECULAR ? 1 : mode == PI_MODE_IRONENERGY ? 8 : 0;
}
/* contribution values per store entry: the losses of IronEnergyMode reach ten
 * tracers (NH4, NO3, DFe and PFe_bio besides the six of the other modes) */
__host__ __device__ constexpr int mode_ncon(int mode) { return mode == PI_MODE_IRONENERGY ? 10 : 6; }
/* values per slot of a reborn daughter's consumption: DIC, DOC, NH4, NO3, PO4, DFe */
#define PI_RCONS 6

/* Pointers read from device memory (the DevState) are generic pointers, and
 * accesses through them compile to FLAT instructions (64-bit addresses, both
 * wait counters).  Kernels access buffers through global-address-space views
 * instead: SGPR-base addressing with 32-bit offsets, vector-memory counter only. */
#define PI_GLOBAL __attribute__((address_space(1)))
template <typename T> __device__ __forceinline__ PI_GLOBAL T* gp(T* p) { return (PI_GLOBAL T*)p; }
#define PI_WAIT_VM0 0x0F70        /* s_waitcnt vmcnt(0): vector memory (incl. loads into LDS) done */

/* Loads and stores through global-address-space pointers: a uniform (SGPR) base plus a
 * 32-bit byte offset, so a load needs no 64-bit address arithmetic and loads at
 * fixed distances share one offset register (immediate offsets) */
__device__ __forceinline__ float ldg(const PI_GLOBAL float* base, uint32_t cell)
{
    return *(const PI_GLOBAL float*)((const PI_GLOBAL char*)base + cell * 4u);
}
__device__ __forceinline__ void stg(PI_GLOBAL float* base, uint32_t cell, float v)
{
    *(PI_GLOBAL float*)((PI_GLOBAL char*)base + cell * 4u) = v;
}

/* two consecutive floats (cells c and c+1) in one 8-byte load (4-byte aligned) */
typedef float pi_f2u __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ pi_f2u ldg2(const PI_GLOBAL float* base, uint32_t cell)
{
    return *(const PI_GLOBAL pi_f2u*)((const PI_GLOBAL char*)base + cell * 4u);
}

struct GStore {
    PI_GLOBAL uint32_t* key;
    PI_GLOBAL float *x, *y, *z, *Sz, *Bm, *Cq, *Nq, *Pq, *Chl, *gen, *age;
    PI_GLOBAL uint8_t* flags;
    PI_GLOBAL uint32_t* cell;
    PI_GLOBAL float* ex[PI_NEX];
};
__device__ __forceinline__ GStore gstore(const Store& s)
{
    GStore g;
    g.key = gp(s.key); g.x = gp(s.x); g.y = gp(s.y); g.z = gp(s.z); g.Sz = gp(s.Sz); g.Bm = gp(s.Bm);
    g.Cq = gp(s.Cq); g.Nq = gp(s.Nq); g.Pq = gp(s.Pq); g.Chl = gp(s.Chl); g.gen = gp(s.gen); g.age = gp(s.age);
    g.flags = gp(s.flags); g.cell = gp(s.cell);
    for (int k = 0; k < PI_NEX; ++k) g.ex[k] = gp(s.ex[k]);
    return g;
}

struct SlotTable {               /* per species, stride cap */
    float *x, *y, *z, *Sz, *Bm, *Cq, *Nq, *Pq, *Chl, *gen, *age;
    uint8_t* flags;
    uint32_t* cell;
    int32_t* death;              /* iteration at which the row was last written */
    float* ex[PI_NEX];
};

struct Contrib {
    float* v[10];                /* v[6..9] only for IronEnergyMode models (else null) */
    uint8_t* tag;
};

struct Vel { const float *u, *v, *w; };

/* 1-based padded (i,j,k) -> linear index */
__device__ __forceinline__ long long gidx(const DevGrid& g, int i, int j, int k) {
    return (long long)(i - 1) + (long long)g.Px * ((long long)(j - 1) + (long long)g.Py * (long long)(k - 1));
}
__device__ __forceinline__ bool LM(const DevGrid& g, long long c) { return g.has_land ? g.lm[c] != 0 : true; }
__device__ __forceinline__ float dzC_at(const DevGrid& g, long long k) {   /* clamp as documented */
    if (k < 1) k = 1;
    if (k > g.Pz) k = g.Pz;
    return g.dzC[k - 1];
}
__device__ __forceinline__ float dzF_at(const DevGrid& g, int k) { return g.dzF[k - 1]; }
/* volume(i, j, k) of the cell c at level k (Grids/areas_volumes_spacings.jl:22,38) */
__device__ __forceinline__ float cell_vol(const DevGrid& g, long long c, int k)
{
    return g.gm ? g.mVol[c] : g.dx * g.dy * dzF_at(g, k);
}
/* column (i, j) of the padded 2-D metric arrays, 1-based */
__device__ __forceinline__ long long col2(const DevGrid& g, int i, int j) { return (long long)(i - 1) + (long long)g.Px * (j - 1); }
/* a / b in 3 operations for the hot paths, given yd = RN_double(1/b): the
 * product rounded once to double then to float, which is RN(a/b) except when
 * a/b is exactly a rounding boundary of the subnormal range (an odd multiple
 * of 2^-150; in the normal range the rounded product is always right, see
 * divc).  A wrong rounding there always yields a subnormal float (the two
 * candidates of an odd multiple M·2^-150, M >= 3, are subnormal; for M = 1 the
 * wrong candidate is 2^-149), so `sub` is raised for every subnormal result and
 * the caller recomputes its work item with divc. */
template <bool EXACT>
__device__ __forceinline__ float divq(float a, float b, double yd, bool& sub)
{
    if (EXACT) return divc(a, b, yd);
    const float f = (float)((double)a * yd);
    sub |= __builtin_isfpclass(f, 0x0090);     /* negative or positive subnormal */
    return f;
}

/* Julia's max / min of Float32 (NaN if either operand is NaN, -0 < +0) are the
 * IEEE 754-2019 maximum / minimum: one instruction each on gfx950.  For NaN
 * operands the NaN may carry another payload than Julia's x - y. */
__device__ __forceinline__ float dmaxf(float x, float y) { return __builtin_elementwise_maximum(x, y); }
__device__ __forceinline__ float dminf(float x, float y) { return __builtin_elementwise_minimum(x, y); }

/* a / dzC_at(g, k) (same clamp) */
__device__ __forceinline__ float div_dzC(const DevGrid& g, float a, long long k) {
    if (k < 1) k = 1;
    if (k > g.Pz) k = g.Pz;
    return divc(a, g.dzC[k - 1], g.ydzC[k - 1]);
}

/* MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin, and
 * each XCD has its own L2.  xcd_block() renumbers the workgroups so that each
 * XCD runs a contiguous range of logical blocks: neighbouring tiles (which
 * share cache lines: the cells of consecutive individuals, the rows j±1 of a
 * tracer stencil) then meet in the same L2.  A bijection for any grid size. */
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, i = b >> 3;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

/* launch helpers */
/* The block a kernel body runs as: its index among the blocks working on one
 * model, and their count.  A single-model launch passes the XCD-renumbered
 * block index and the grid size; a batched launch looks its block up in a
 * table of (model, index, count), so one launch steps many models. */
struct Blk { uint32_t b, n; };

/* The two entry points of a kernel body BODY(const ARGS&, Blk): NAME for one
 * model (arguments by value; BIDX = blockIdx.x, or xcd_block() for kernels
 * whose neighbouring blocks share cache lines) and NAME##_m for a batch
 * (argument table + block table). */
#define PI_KERNEL3(NAME, BODY, ARGS, THREADS, BIDX)                                                  \
    __global__ void __launch_bounds__(THREADS) NAME(ARGS a) { BODY(a, Blk{BIDX, gridDim.x}); }      \
    __global__ void __launch_bounds__(THREADS) NAME##_m(const ARGS* __restrict__ A,                  \
                                                        const uint4* __restrict__ bt)                 \
    {                                                                                                 \
        const uint4 e = bt[blockIdx.x];                                                               \
        BODY(A[e.x], Blk{e.y, e.z});                                                                  \
    }
/* PI_KERNEL3 with at least WAVES waves per SIMD (the register budget the
 * compiler must fit: 512 / WAVES VGPRs) */
#define PI_KERNEL3W(NAME, BODY, ARGS, THREADS, BIDX, WAVES)                                              \
    __global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WAVES)))             \
    NAME(ARGS a) { BODY(a, Blk{BIDX, gridDim.x}); }                                                     \
    __global__ void __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(WAVES)))             \
    NAME##_m(const ARGS* __restrict__ A, const uint4* __restrict__ bt)                                  \
    {                                                                                                   \
        const uint4 e = bt[blockIdx.x];                                                                 \
        BODY(A[e.x], Blk{e.y, e.z});                                                                    \
    }
#define PI_KERNEL2(NAME, BODY, ARGS, THREADS) PI_KERNEL3(NAME, BODY, ARGS, THREADS, xcd_block())
#define PI_KERNEL(NAME, ARGS, THREADS) PI_KERNEL3(NAME, NAME##_body, ARGS, THREADS, blockIdx.x)
#define PI_KERNEL_XCD(NAME, ARGS, THREADS) PI_KERNEL3(NAME, NAME##_body, ARGS, THREADS, xcd_block())

static inline int grid_for(long long n, int block, int maxblocks) {
    long long b = (n + block - 1) / block;
    if (b > maxblocks) b = maxblocks;
    if (b < 1) b = 1;
    return (int)b;
}

#define PI_MAXSP 31
/* The queue of entries that change cell in a step is split in PI_NQ shards,
 * shard (e / 256) mod PI_NQ: one device-wide atomic counter per shard (a
 * single counter serialises every wave's append: ~11 ns each on MI355X). */
#define PI_NQ 64

/* Everything a step flips or counts lives on the device, so a step is a fixed
 * sequence of launches with fixed arguments (capturable in a graph) and an
 * error raised early in a step leaves every buffer untouched. */
struct DevState {
    Store store[2];                    /* live individuals: [scur] current, [1-scur] next */
    float* tr[2][PI_NTRACERS];         /* tracers: [tcur] current */
    float* par[2];                     /* [pcur] = par of the previous step (par₀) */
    int32_t scur, tcur, pcur, vcur;
    long long iteration;               /* completed steps */
    unsigned long long ind_steps;      /* Σ over completed steps of the individuals stepped */
    uint32_t nstore;                   /* entries in store[scur] */
    uint32_t nA;                       /* of which the first nA are sorted by (cell, key) as laid out by the previous seg */
    uint32_t nsmall[PI_NQ];            /* queued entries of this step, per queue shard */
    uint32_t nmix[PI_NQ];              /* re-sorted cells of this step, per shard (of the queue position) */
    uint32_t err, err_sp;
    uint32_t nbig;
    uint32_t nlong;                    /* cells longer than LONG_SEG this step */
    uint32_t dtot[PI_MAXSP], ftot[PI_MAXSP], doff[PI_MAXSP];
};
