/*
 * k_util.hip - device-wide exclusive scan, capacity check, end-of-step flip,
 * and the slot-table <-> store conversions used by uploads and downloads.
 */
#include "pi_device.h"
#include "k_args.h"

#define SCAN_THREADS 256
#define SCAN_ITEMS 16
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

/* Input of a scan: either a plain u32 array, or the division bitmaps laid out as
 * [popcount(D words) of all species | popcount(free words) of all species]. */

__device__ __forceinline__ uint32_t scan_len(const ScanIn& s) { return s.dn ? min(*s.dn, s.n) : s.n; }

__device__ __forceinline__ uint32_t scan_load(const ScanIn& s, uint32_t i, uint32_t n)
{
    if (i >= n) return 0;
    if (!s.bits) return s.a[i];
    const uint32_t half = (uint32_t)s.nsp * (uint32_t)s.nw;
    if (i < half) return __popc(s.D[i]);
    const uint32_t j = i - half, w = j % (uint32_t)s.nw;
    const long long rem = s.cap - (long long)w * 32;
    const uint32_t valid = rem >= 32 ? 0xFFFFFFFFu : ((1u << rem) - 1u);
    return __popc(~s.A[j] & valid);
}

/* the SCAN_ITEMS consecutive inputs of a thread: 16-byte loads when they are
 * whole (thread bases are multiples of SCAN_ITEMS, a multiple of 4) */
__device__ __forceinline__ void scan_loadv(const ScanIn& s, uint32_t base, uint32_t n, uint32_t* v)
{
    if (!s.bits && base + SCAN_ITEMS <= n) {
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k += 4) {
            const uint4 a = *(const uint4*)(s.a + base + k);
            v[k] = a.x; v[k + 1] = a.y; v[k + 2] = a.z; v[k + 3] = a.w;
        }
    } else {
        for (int k = 0; k < SCAN_ITEMS; ++k) v[k] = scan_load(s, base + k, n);
    }
}

/* exclusive scan of one value per thread across a 256-thread block */
__device__ __forceinline__ uint32_t block_exscan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        if (w < wid) off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

__device__ __forceinline__ void k_scan_partials_impl(Blk K, ScanIn s, uint32_t* partials)
{
    const uint32_t n = scan_len(s);
    if (K.b * SCAN_TILE >= n) {               /* beyond a device-side length */
        if (threadIdx.x == 0) partials[K.b] = 0;
        return;
    }
    const uint32_t base = K.b * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], sum = 0;
    scan_loadv(s, base, n, v);
    for (int k = 0; k < SCAN_ITEMS; ++k) sum += v[k];
    uint32_t tot;
    block_exscan(sum, &tot);
    if (threadIdx.x == 0) partials[K.b] = tot;
}
__device__ __forceinline__ void k_scan_partials_body(const ScanArgs& x, Blk K) { k_scan_partials_impl(K, x.in, x.partials); }
PI_KERNEL(k_scan_partials, ScanArgs, 256)

/* Tile b's exclusive prefix (the sum of the partials of tiles 0..b-1, a block
 * reduction over at most a few thousand values), then its outputs; the last
 * tile also writes the total to out[n].  One kernel where a single-block scan
 * of the partials would be a launch of its own. */
__device__ __forceinline__ void k_scan_final_impl(Blk K, ScanIn s, const uint32_t* partials, uint32_t* out, uint32_t nb)
{
    __shared__ uint32_t s_pre;
    uint32_t acc = 0;
    for (uint32_t i = threadIdx.x * 4u; i < K.b; i += blockDim.x * 4u) {
        if (i + 4u <= K.b) {
            const uint4 p = *(const uint4*)(partials + i);                  /* i is a multiple of 4 */
            acc += p.x + p.y + p.z + p.w;
        } else {
            for (uint32_t k = i; k < K.b; ++k) acc += partials[k];
        }
    }
    uint32_t atot;
    block_exscan(acc, &atot);
    if (threadIdx.x == 0) s_pre = atot;
    __syncthreads();
    const uint32_t n = scan_len(s);
    const uint32_t base = K.b * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], sum = 0;
    scan_loadv(s, base, n, v);
    for (int k = 0; k < SCAN_ITEMS; ++k) sum += v[k];
    uint32_t tot;
    uint32_t pre = block_exscan(sum, &tot) + s_pre;
    if (K.b == nb - 1 && threadIdx.x == 0) out[s.n] = s_pre + tot;     /* values past n are 0 */
    if (K.b * SCAN_TILE > n) return;
    uint32_t o[SCAN_ITEMS];
    for (int k = 0; k < SCAN_ITEMS; ++k) { o[k] = pre; pre += v[k]; }
    if (base + SCAN_ITEMS <= n && base + SCAN_ITEMS <= s.n) {        /* outputs are 16-byte aligned */
#pragma unroll
        for (int k = 0; k < SCAN_ITEMS; k += 4) *(uint4*)(out + base + k) = make_uint4(o[k], o[k + 1], o[k + 2], o[k + 3]);
    } else {
        for (int k = 0; k < SCAN_ITEMS; ++k)
            if (base + k <= n && base + k < s.n) out[base + k] = o[k];   /* out[n] = total when n < s.n */
    }
}
__device__ __forceinline__ void k_scan_final_body(const ScanArgs& x, Blk K) { k_scan_final_impl(K, x.in, x.partials, x.out, x.nb); }
PI_KERNEL(k_scan_final, ScanArgs, 256)

/* totals, daughter offsets and the capacity check (plankton_update.jl:18-22:
 * dot(dvid, ac) > length(findall(!ac)) throws). */
__device__ __forceinline__ void k_capcheck_impl(Blk K, DevState* st, const uint32_t* G, int nsp, int nw)
{
    if (threadIdx.x != 0 || st->err) return;
    uint32_t off = 0;
    for (int s = 0; s < nsp; ++s) {
        const uint32_t d = G[(uint32_t)(s + 1) * nw] - G[(uint32_t)s * nw];
        const uint32_t f = G[(uint32_t)(nsp + s + 1) * nw] - G[(uint32_t)(nsp + s) * nw];
        st->dtot[s] = d;
        st->ftot[s] = f;
        st->doff[s] = off;
        off += d;
        if (d > f && !st->err) { st->err = PI_ERR_CAPACITY; st->err_sp = (uint32_t)s; }
    }
    st->nbig = 0;
    st->nlong = 0;
    for (int q = 0; q < PI_NQ; ++q) { st->nsmall[q] = 0; st->nmix[q] = 0; }
}
__device__ __forceinline__ void k_capcheck_body(const CapArgs& x, Blk K) { k_capcheck_impl(K, x.st, x.G, x.nsp, x.nw); }
PI_KERNEL(k_capcheck, CapArgs, 64)

/* end of TimeStep!: vel₀ <- vel₁, par₀ <- par, new tracers and store become current */
__device__ __forceinline__ void k_finish_impl(Blk K, DevState* st, const uint32_t* seg, long long ncell, int nsp)
{
    if (threadIdx.x != 0 || st->err) return;
    uint32_t born = 0;
    for (int s = 0; s < nsp; ++s) born += st->dtot[s];
    st->nA = seg[ncell];                      /* the sorted part of the next store */
    st->ind_steps += seg[ncell];              /* individuals stepped (alive at this step's gather) */
    st->nstore = seg[ncell] + born;
    st->scur ^= 1;
    st->tcur ^= 1;
    st->pcur ^= 1;
    st->iteration += 1;
}
__device__ __forceinline__ void k_finish_body(const FinishArgs& x, Blk K) { k_finish_impl(K, x.st, x.seg, x.ncell, x.nsp); }
PI_KERNEL(k_finish, FinishArgs, 64)


/* a plain streaming copy (16 bytes per lane per iteration) */
__global__ void __launch_bounds__(256) k_probe_copy(uint4* dst, const uint4* src, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_zero_u32(uint32_t* p, long long n, const DevState* st)
{
    if (st && st->err) return;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}

/* vel½ = (vel₀ + vel₁) * 0.5f0 (time_step.jl:17-19), u, v and w */
__device__ __forceinline__ void k_velhalf_body(const VelhalfArgs& x, Blk K)
{
    if (x.st->err) return;
    for (int c = 0; c < 3; ++c) {
        const float* a = x.a[c];
        const float* b = x.b[c];
        float* h = x.h[c];
        for (long long i = K.b * (long long)blockDim.x + threadIdx.x; i < x.n; i += (long long)K.n * blockDim.x)
            h[i] = (a[i] + b[i]) * 0.5f;
    }
}
PI_KERNEL(k_velhalf, VelhalfArgs, 256)

/* diags.tracer[key] .+= tracers[key].data over padded arrays (time_step.jl:168-172) */
__device__ __forceinline__ void k_accum_body(const AccumArgs& x, Blk K)
{
    const DevState* st = x.st;
    if (st->err) return;
    for (int j = 0; j < x.nq; ++j) {
        const float* src = st->tr[st->tcur ^ 1][x.q[j]];   /* the tracers written by this step */
        float* d = x.d[j];
        for (long long i = K.b * (long long)blockDim.x + threadIdx.x; i < x.n; i += (long long)K.n * blockDim.x)
            d[i] = d[i] + src[i];
    }
}
PI_KERNEL(k_accum, AccumArgs, 256)

/* ---------------------------------------------------- store <-> slot table */
/* live store entries -> their slot-table rows (marks them active there) */
__global__ void k_store_to_table(DevState* st, SlotTable tab, long long cap)
{
    const Store S = st->store[st->scur];
    const uint32_t n = st->nstore;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const uint8_t fl = S.flags[e];
        if (!(fl & F_AC)) continue;
        const uint32_t key = S.key[e];
        const long long r = (long long)(key >> PI_SLOT_BITS) * cap + (key & PI_SLOT_MASK);
        tab.x[r] = S.x[e]; tab.y[r] = S.y[e]; tab.z[r] = S.z[e];
        tab.Sz[r] = S.Sz[e]; tab.Bm[r] = S.Bm[e]; tab.Cq[r] = S.Cq[e]; tab.Nq[r] = S.Nq[e];
        tab.Pq[r] = S.Pq[e]; tab.Chl[r] = S.Chl[e]; tab.gen[r] = S.gen[e]; tab.age[r] = S.age[e];
        for (int k = 0; k < PI_NEX; ++k)
            if (tab.ex[k]) tab.ex[k][r] = S.ex[k][e];
        tab.flags[r] = fl;
        tab.cell[r] = S.cell[e];
        tab.death[r] = (int32_t)st->iteration;
    }
}

/* slot-table rows marked active -> a fresh store (order is irrelevant: the
 * next step sorts by cell) */
__global__ void k_rebuild(DevState* st, SlotTable tab, long long cap, int nsp, uint8_t* sflag)
{
    const Store S = st->store[st->scur];
    const long long total = cap * nsp;
    for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < total; r += (long long)gridDim.x * blockDim.x) {
        const uint8_t fl = tab.flags[r];
        sflag[r] = (fl & F_AC) ? (uint8_t)(fl & (F_AC | F_GRAZ | F_MORT | F_DVID)) : (uint8_t)0;
        if (!(fl & F_AC)) continue;
        const uint32_t e = atomicAdd(&st->nstore, 1u);
        S.key[e] = ((uint32_t)(r / cap) << PI_SLOT_BITS) | (uint32_t)(r % cap);
        S.x[e] = tab.x[r]; S.y[e] = tab.y[r]; S.z[e] = tab.z[r];
        S.Sz[e] = tab.Sz[r]; S.Bm[e] = tab.Bm[r]; S.Cq[e] = tab.Cq[r]; S.Nq[e] = tab.Nq[r];
        S.Pq[e] = tab.Pq[r]; S.Chl[e] = tab.Chl[r]; S.gen[e] = tab.gen[r]; S.age[e] = tab.age[r];
        for (int k = 0; k < PI_NEX; ++k)
            if (tab.ex[k]) S.ex[k][e] = tab.ex[k][r];
        S.flags[e] = fl;
        S.cell[e] = tab.cell[r];
    }
}

/* One column of `model.individuals.phytos.spN.data` in the reference layout.
 * Rows of inactive slots written before the current iteration have gone
 * through at least one step while inactive, whose fixed point is: x = y = z = 0
 * (boundaries * ac), xi = yi = zi = H (find_inds), Sz = Bm / (Cquota * Nsuper)
 * (update_cellsize! runs on every slot; QuotaMode and CarbonMode) and zero
 * decisions; every other field is unchanged by a step with ac = false, except
 * in MacroMolecularMode, whose inactive rows k_idlem steps every step. */
__global__ void k_table_field(const DevState* st, SlotTable tab, int sp, long long cap, int field,
                              const float* params, DevGrid g, void* out, int mode)
{
    const bool mm = mode == PI_MODE_MACROMOLECULAR || mode == PI_MODE_IRONENERGY;   /* k_idle* steps their rows */
    const long long it = st->iteration;
    const float* q = params + sp * PI_NPARAMS;
    for (long long s = blockIdx.x * (long long)blockDim.x + threadIdx.x; s < cap; s += (long long)gridDim.x * blockDim.x) {
        const long long r = (long long)sp * cap + s;
        const uint8_t fl = tab.flags[r];
        const bool act = (fl & F_AC) != 0;
        const bool settled = !act && (long long)tab.death[r] < it;
        float v = 0.0f;
        switch (field) {
            case PI_F_x: v = settled ? 0.0f : tab.x[r]; break;
            case PI_F_y: v = settled ? 0.0f : tab.y[r]; break;
            case PI_F_z: v = settled ? 0.0f : tab.z[r]; break;
            case PI_F_Sz: v = (settled && !mm) ? tab.Bm[r] / (q[PI_P_Cquota] * q[PI_P_Nsuper]) : tab.Sz[r]; break;
            case PI_F_Bm: v = tab.Bm[r]; break;
            case PI_F_Cq: v = tab.Cq[r]; break;
            case PI_F_Nq: v = tab.Nq[r]; break;
            case PI_F_Pq: v = tab.Pq[r]; break;
            case PI_F_Chl: v = tab.Chl[r]; break;
            case PI_F_gen: v = tab.gen[r]; break;
            case PI_F_age: v = tab.age[r]; break;
            case PI_F_graz: v = (!settled && (fl & F_GRAZ)) ? 1.0f : 0.0f; break;
            case PI_F_mort: v = (!settled && (fl & F_MORT)) ? 1.0f : 0.0f; break;
            case PI_F_dvid: v = (!settled && (fl & F_DVID)) ? 1.0f : 0.0f; break;
            default:
                if (field >= PI_F_ex0 && tab.ex[field - PI_F_ex0]) v = tab.ex[field - PI_F_ex0][r];
                break;
        }
        if (field <= PI_F_dvid || field >= PI_F_ex0) { ((float*)out)[s] = v; continue; }
        if (field == PI_F_ac) { ((uint8_t*)out)[s] = act ? 1 : 0; continue; }
        long long xi, yi, zi;
        if (settled) { xi = g.Hx; yi = g.Hy; zi = g.Hz; }
        else if (fl & F_ZERO) { xi = yi = zi = 0; }
        else {
            const uint32_t c = tab.cell[r];
            xi = c % (uint32_t)g.Px + 1;
            yi = (c / (uint32_t)g.Px) % (uint32_t)g.Py + 1;
            zi = c / ((uint32_t)g.Px * (uint32_t)g.Py) + 1;
        }
        ((long long*)out)[s] = field == PI_F_xi ? xi : field == PI_F_yi ? yi : zi;
    }
}

__global__ void k_count(const DevState* st, int sp, unsigned long long* out)
{
    const Store S = st->store[st->scur];
    const uint32_t n = st->nstore;
    unsigned long long c = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        if ((S.flags[e] & F_AC) && (S.key[e] >> PI_SLOT_BITS) == (uint32_t)sp) ++c;
    if (c) atomicAdd(out, c);
}

/* -------------------------------------------------- initial state on device
 * initialize_plankton! (QuotaMode/plankton_generation.jl:35-72; CarbonMode and
 * MacroMolecularMode alike, plankton_generation.jl of each) for every slot
 * of every species, written into the slot table like an upload of the host
 * arrays (ModelSpec in model.py computes the same values on the host with the
 * same stream and the same binary64 log/cos). */
__global__ void __launch_bounds__(256) k_init_individuals(InitArgs a)
{
    const DevGrid& g = a.g;
    const long long total = a.cap * a.nsp;
    const int32_t death = (int32_t)a.st->iteration;
    for (long long r = blockIdx.x * (long long)blockDim.x + threadIdx.x; r < total; r += (long long)gridDim.x * blockDim.x) {
        const int sp = (int)(r / a.cap);
        const long long i = r - (long long)sp * a.cap;
        const float* q = a.params + sp * PI_NPARAMS;
        float ux, uy, uz, uc, un, up, ua, ub, u3;
        pi_uniform3(a.seed, 0, sp, PI_KIND_INIT_POS, (uint64_t)i, &ux, &uy, &uz);
        pi_uniform3(a.seed, 0, sp, PI_KIND_INIT_QUOTA, (uint64_t)i, &uc, &un, &up);
        pi_uniform3(a.seed, 0, sp, PI_KIND_INIT_SIZE, (uint64_t)i, &ua, &ub, &u3);
        bool ac = i < a.N[sp];
        const float x = jl_maskf(ac, ux * (float)g.Nx);
        const float y = jl_maskf(ac, uy * (float)g.Ny);
        const float z = jl_maskf(ac, uz * (float)g.Nz);
        float Sz, Bm, Cq, Nq, Pq, Chl, RNA = 0.0f;
        float fe[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};   /* IronEnergyMode: qNO3, qFe, qFePS, qFeNR, qFeNF */
        if (a.mode == PI_MODE_IRONENERGY) {
            /* IronEnergyMode/plankton_generation.jl:48-94: the size normal as QuotaMode's,
             * CH, qNO3, qNH4 from the quota triple, qP from the size triple's third
             * draw, qFe from its own triple; the maxima at a tenth; the iron split
             * into photosynthesis, nitrate reduction and N fixation pools */
            float uf, uf1, uf2;
            pi_uniform3(a.seed, 0, sp, PI_KIND_INIT_FE, (uint64_t)i, &uf, &uf1, &uf2);
            const float Sz0 = pi_init_normal(ua, ub);
            Sz = jl_maskf(ac, jl_maxf(1.0f, Sz0 * q[PI_I_var] + q[PI_I_mean]));
            Bm = q[PI_I_Cquota] * Sz * q[PI_I_Nsuper];
            Cq = uc * (q[PI_I_CHmax] * 0.1f) * Bm;                            /* CH */
            fe[0] = un * ((q[PI_I_qNO3max] * 0.1f) * (Bm + Cq));              /* qNO3 */
            Nq = up * ((q[PI_I_qNH4max] * 0.1f) * (Bm + Cq));                 /* qNH4 */
            Pq = u3 * ((q[PI_I_qPmax] * 0.1f) * (Bm + Cq));                   /* qP */
            float qFe = uf * ((q[PI_I_qFemax] * 0.1f) * (Bm + Cq));
            fe[2] = qFe * 0.4f;
            fe[3] = qFe * 0.3f * q[PI_I_is_nr];
            fe[4] = qFe * 0.3f * (q[PI_I_is_croc] + q[PI_I_is_tric]);
            fe[1] = qFe - (fe[2] + fe[3] + fe[4]);
            Chl = Bm * q[PI_I_Chl2Cint];
        } else if (a.mode == PI_MODE_MACROMOLECULAR) {
            /* MacroMolecularMode/plankton_generation.jl:31-78: DNA, RNA and protein from
             * one draw triple, CH, NST and PST from another; Sz = DNA, Bm = PRO,
             * Cq = CH, Nq = NST, Pq = PST */
            const float var = q[PI_M_var], cN = q[PI_M_C_DNA], Ns = q[PI_M_Nsuper];
            const float DNA = jl_maskf(ac, (ua * var + 1.0f) * cN * Ns);
            RNA = jl_maskf(ac, (ub * var * 2.0f + 1.0f - var) * cN * Ns) * q[PI_M_RNA2DNA];
            Bm = jl_maskf(ac, (u3 * var * 2.0f + 1.0f - var) * cN * Ns) * q[PI_M_PRO2DNA];
            Cq = (uc * var * 2.0f + 1.0f - var) * DNA * q[PI_M_CH2DNA];
            Nq = (un * var * 2.0f + 1.0f - var) * Cq * 16.0f / 106.0f;
            Pq = (up * var * 2.0f + 1.0f - var) * Cq / 106.0f;
            Chl = DNA * q[PI_M_Chl2DNA] * 893.49f / 55.0f;
            Sz = DNA;
        } else {
            const float Sz0 = pi_init_normal(ua, ub);
            Sz = jl_maskf(ac, jl_maxf(1.0f, Sz0 * q[PI_P_var] + q[PI_P_mean]));
            Bm = (q[PI_P_Cquota] * Sz) * q[PI_P_Nsuper];     /* PI_C_Cquota, PI_C_Nsuper alike */
            const bool carbon = a.mode == PI_MODE_CARBON;   /* CarbonMode/plankton_generation.jl:34-56: Bd = 0 */
            Cq = carbon ? 0.0f : (uc * q[PI_P_Cqmax]) * Bm;
            Nq = carbon ? 0.0f : un * (q[PI_P_Nqmax] * (Bm + Cq) - Bm * q[PI_P_R_NC]);
            Pq = carbon ? 0.0f : up * (q[PI_P_Pqmax] * (Bm + Cq) - Bm * q[PI_P_R_PC]);
            Chl = Bm * q[carbon ? (int)PI_C_Chl2C : (int)PI_P_Chl2Cint];
        }
        if (ac && g.has_land) {   /* mask_individuals! (Individuals/utils.jl:89-100) */
            const int xi = (int)truncf(x + 1.0f) + g.Hx, yi = (int)truncf(y + 1.0f) + g.Hy, zi = (int)truncf(z + 1.0f) + g.Hz;
            ac = g.lm[gidx(g, xi, yi, zi)] != 0;
        }
        a.tab.x[r] = x; a.tab.y[r] = y; a.tab.z[r] = z;
        a.tab.Sz[r] = Sz; a.tab.Bm[r] = Bm; a.tab.Cq[r] = Cq; a.tab.Nq[r] = Nq; a.tab.Pq[r] = Pq;
        a.tab.Chl[r] = Chl;
        if (a.mode == PI_MODE_MACROMOLECULAR) a.tab.ex[0][r] = RNA;
        if (a.mode == PI_MODE_IRONENERGY) {
            for (int k = 0; k < 5; ++k) a.tab.ex[k][r] = fe[k];
            a.tab.ex[5][r] = a.tab.ex[6][r] = a.tab.ex[7][r] = 0.0f;     /* tdark, exEn, ptc */
        }
        a.tab.gen[r] = i < a.N[sp] ? 1.0f : 0.0f;
        a.tab.age[r] = 0.0f;
        a.tab.flags[r] = (uint8_t)((ac ? F_AC : 0) | F_ZERO);
        a.tab.cell[r] = 0u;
        a.tab.death[r] = death;
    }
}

/* generate_tracers with the values source: c0 · vals[pick] per padded cell,
 * vals = lower : step : upper, pick from the tracer's sub-stream; then the
 * landmask (Bool mask keeps the sign of zero) */
__global__ void __launch_bounds__(256) k_init_tracers(TracerInitArgs a)
{
    const DevGrid& g = a.g;
    const long long n = (long long)g.ncell * PI_NTRACERS;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t / g.ncell);
        const long long cell = t - (long long)q * g.ncell;
        float u, u1, u2;
        pi_uniform3(a.seed, 0, 0, PI_KIND_TRACER + q, (uint64_t)cell, &u, &u1, &u2);
        long long pick = (long long)((double)u * (double)a.n[q]);
        if (pick > a.n[q] - 1) pick = a.n[q] - 1;
        const float v = (float)((double)a.lower[q] + (double)pick * (double)a.step);
        const float c = a.c0[q] * v;
        a.slab[t] = (g.has_land && !g.lm[cell]) ? copysignf(0.0f, c) : c;
    }
}

void launch_init_individuals(hipStream_t s, const InitArgs& a)
{
    hipLaunchKernelGGL(k_init_individuals, dim3(grid_for(a.cap * a.nsp, 256, 8192)), dim3(256), 0, s, a);
}
void launch_init_tracers(hipStream_t s, const TracerInitArgs& a)
{
    hipLaunchKernelGGL(k_init_tracers, dim3(grid_for((long long)a.g.ncell * PI_NTRACERS, 256, 8192)), dim3(256), 0, s, a);
}

/* ------------------------------------------------------------- registry */
uint32_t scan_tiles(uint32_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

void reg_util(KInfo* k)
{
    PI_REG(k, K_VELHALF, k_velhalf, VelhalfArgs, 256, true);
    PI_REG(k, K_SCAN_P, k_scan_partials, ScanArgs, SCAN_THREADS, false);
    PI_REG(k, K_SCAN_F, k_scan_final, ScanArgs, SCAN_THREADS, false);
    PI_REG(k, K_CAPCHECK, k_capcheck, CapArgs, 64, false);
    PI_REG(k, K_ACCUM, k_accum, AccumArgs, 256, true);
    PI_REG(k, K_FINISH, k_finish, FinishArgs, 64, false);
}

/* ------------------------------------------------------------- launchers */
ScanIn scan_plain(const uint32_t* a, uint32_t n)
{
    ScanIn s{}; s.a = a; s.n = n; s.bits = 0; return s;
}
ScanIn scan_bits(const uint32_t* D, const uint32_t* A, int nsp, int nw, long long cap)
{
    ScanIn s{}; s.bits = 1; s.D = D; s.A = A; s.nsp = nsp; s.nw = nw; s.cap = cap; s.n = 2u * nsp * nw; return s;
}
void launch_zero(hipStream_t s, uint32_t* p, long long n, const DevState* st)
{
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, p, n, st);
}
void launch_store_to_table(hipStream_t s, DevState* st, const SlotTable& tab, long long cap, long long maxn)
{
    hipLaunchKernelGGL(k_store_to_table, dim3(grid_for(maxn, 256, 8192)), dim3(256), 0, s, st, tab, cap);
}
void launch_rebuild(hipStream_t s, DevState* st, const SlotTable& tab, long long cap, int nsp, uint8_t* sflag)
{
    hipLaunchKernelGGL(k_rebuild, dim3(grid_for(cap * nsp, 256, 8192)), dim3(256), 0, s, st, tab, cap, nsp, sflag);
}
void launch_probe_copy(hipStream_t s, void* dst, const void* src, long long n16)
{
    hipLaunchKernelGGL(k_probe_copy, dim3(grid_for(n16, 256, 8192)), dim3(256), 0, s, (uint4*)dst, (const uint4*)src, n16);
}
void launch_table_field(hipStream_t s, const DevState* st, const SlotTable& tab, int sp, long long cap, int field,
                        const float* params, const DevGrid& g, void* out, int mode)
{
    hipLaunchKernelGGL(k_table_field, dim3(grid_for(cap, 256, 8192)), dim3(256), 0, s, st, tab, sp, cap, field, params, g,
                       out, mode);
}
void launch_count(hipStream_t s, const DevState* st, int sp, unsigned long long* out, long long maxn)
{
    hipLaunchKernelGGL(k_count, dim3(grid_for(maxn, 256, 2048)), dim3(256), 0, s, st, sp, out);
}
