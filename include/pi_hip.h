/*
 * pi_hip.h - C ABI of libpi_hip.so, the MI355X (gfx950) implementation of the
 * PlanktonIndividuals.jl per-timestep hot path (QuotaMode physiology, RK4
 * particle advection, random-walk diffusion, division/death, particle->grid
 * consumption, DST3FL tracer advection + diffusion + sinking + forcing).
 *
 * The ABI is what a new `HIPArch <: Architecture` (src/Architectures.jl:16)
 * binds with `ccall`.  Plain C types only: pointers, sizes, POD structs.
 *
 * Conventions
 *  - Arrays are column-major (x fastest), exactly like the reference's Julia
 *    arrays.  "Padded" fields have shape (Nx+2Hx, Ny+2Hy, Nz+2Hz)
 *    (src/Biogeochemistry/tracer_fields.jl:10-14).
 *  - Individuals are per species, per slot, length `capacity`
 *    (`max_individuals`, src/Model/models.jl:54), slot order = the
 *    reference's StructArray index order (QuotaMode/plankton_generation.jl:2-13).
 *  - Host pointers are borrowed for the duration of the call only.  The
 *    library owns every device buffer.  A handle is not re-entrant; calls on
 *    one handle are ordered on the handle's own HIP stream.
 *  - Every function returns a pi_status; pi_last_error() gives the text.
 */
#ifndef PI_HIP_H
#define PI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (errors are never thrown across the ABI) ---------------- */
typedef enum pi_status {
    PI_OK = 0,
    PI_ERR_ARG = 1,        /* bad argument: mirrors the reference's ArgumentError from validators */
    PI_ERR_CAPACITY = 2,   /* "number of individual exceeds the capacity" (QuotaMode/plankton_update.jl:20-22) */
    PI_ERR_HIP = 3,        /* a HIP runtime call failed */
    PI_ERR_UNSUPPORTED = 4,/* feature outside the implemented hot path (e.g. other physiology modes) */
    PI_ERR_NOMEM = 5
} pi_status;

/* ---- topology (src/Grids/Grids.jl: Periodic / Bounded) -------------------- */
enum { PI_PERIODIC = 0, PI_BOUNDED = 1 };

/* ---- tracer ids, in the reference's tracer_names order (Biogeochemistry.jl:20) */
enum {
    PI_DIC = 0, PI_NH4, PI_NO3, PI_PO4, PI_DFe, PI_DOC, PI_DON, PI_DOP,
    PI_PFe_inorg, PI_POC, PI_PON, PI_POP, PI_PFe_bio, PI_Dust,
    PI_NTRACERS = 14
};

/* ---- QuotaMode parameters: the first 40 of a species' PI_NPARAMS floats, in
 *      the reference's `param_names` order (QuotaMode/plankton_generation.jl:16-20);
 *      every mode's row is PI_NPARAMS floats, the unused tail zero. ------- */
enum {
    PI_P_Nsuper = 0, PI_P_Cquota, PI_P_mean, PI_P_var, PI_P_Chl2Cint, PI_P_alpha,
    PI_P_Phi, PI_P_Topt, PI_P_Tmax, PI_P_Ea, PI_P_PCmax, PI_P_VDOCmax, PI_P_VNO3max,
    PI_P_VNH4max, PI_P_VPO4max, PI_P_KsatDOC, PI_P_KsatNH4, PI_P_KsatNO3, PI_P_KsatPO4,
    PI_P_Cqmax, PI_P_Nqmax, PI_P_Pqmax, PI_P_Chl2N, PI_P_R_NC, PI_P_R_PC, PI_P_k_mtb,
    PI_P_respir, PI_P_grz_P, PI_P_dvid_type, PI_P_dvid_P, PI_P_dvid_reg, PI_P_dvid_reg2,
    PI_P_mort_P, PI_P_mort_reg, PI_P_grazFracC, PI_P_grazFracN, PI_P_grazFracP,
    PI_P_mortFracC, PI_P_mortFracN, PI_P_mortFracP,
    PI_NQPARAMS = 40,
    PI_NPARAMS = 64
};

/* ---- physiology modes (src/PlanktonIndividuals.jl:65-81: CarbonMode, QuotaMode,
 *      MacroMolecularMode, IronEnergyMode) */
enum { PI_MODE_QUOTA = 0, PI_MODE_CARBON = 1, PI_MODE_MACROMOLECULAR = 2, PI_MODE_IRONENERGY = 3 };

/* ---- CarbonMode parameters: 24 floats in the reference's `param_names` order
 *      (CarbonMode/plankton_generation.jl:14-16), in the first 24 of a species'
 *      PI_NPARAMS floats (the rest unused). ------------------------------- */
enum {
    PI_C_Nsuper = 0, PI_C_Cquota, PI_C_mean, PI_C_var, PI_C_alpha, PI_C_Phi, PI_C_Topt, PI_C_Tmax,
    PI_C_Ea, PI_C_PCmax, PI_C_Chl2C, PI_C_respir, PI_C_f_T2B, PI_C_grz_P, PI_C_dvid_type, PI_C_dvid_P,
    PI_C_dvid_reg, PI_C_dvid_reg2, PI_C_mort_P, PI_C_mort_reg, PI_C_grazFracC, PI_C_mortFracC,
    PI_C_thermal, PI_C_is_bact,
    PI_NCPARAMS = 24
};

/* ---- MacroMolecularMode parameters: 47 floats in the reference's `param_names`
 *      order (MacroMolecularMode/plankton_generation.jl:15-21). ----------- */
enum {
    PI_M_Nsuper = 0, PI_M_C_DNA, PI_M_var, PI_M_PRO2DNA, PI_M_RNA2DNA, PI_M_CH2DNA, PI_M_Chl2DNA,
    PI_M_alpha, PI_M_Phi, PI_M_Topt, PI_M_Tmax, PI_M_Ea,
    PI_M_PCmax, PI_M_VDOCmax, PI_M_VNO3max, PI_M_VNH4max, PI_M_VPO4max, PI_M_KsatDOC, PI_M_KsatNH4,
    PI_M_KsatNO3, PI_M_KsatPO4, PI_M_CHmax, PI_M_NSTmax, PI_M_PSTmax,
    PI_M_Chl2N, PI_M_R_NC_PRO, PI_M_R_NC_DNA, PI_M_R_NC_RNA, PI_M_R_PC_DNA, PI_M_R_PC_RNA, PI_M_respir,
    PI_M_k_pro, PI_M_k_sat_pro, PI_M_k_rna, PI_M_k_sat_rna, PI_M_k_dna, PI_M_k_sat_dna,
    PI_M_dvid_P, PI_M_grz_P, PI_M_mort_P, PI_M_mort_reg, PI_M_grazFracC, PI_M_grazFracN, PI_M_grazFracP,
    PI_M_mortFracC, PI_M_mortFracN, PI_M_mortFracP,
    PI_NMPARAMS = 47
};

/* ---- IronEnergyMode parameters: 64 floats in the reference's `param_names`
 *      order (IronEnergyMode/plankton_generation.jl:28-40). ---------------- */
enum {
    PI_I_Nsuper = 0, PI_I_Cquota, PI_I_SA, PI_I_mean, PI_I_var, PI_I_Chl2Cint,
    PI_I_alpha, PI_I_Topt, PI_I_Tmax, PI_I_Ea, PI_I_is_nr, PI_I_is_croc, PI_I_is_tric,
    PI_I_PCmax, PI_I_VNO3max, PI_I_VNH4max, PI_I_VPO4max,
    PI_I_k_cf, PI_I_k_rs, PI_I_k_nr, PI_I_k_nf, PI_I_k_mtb,
    PI_I_e_cf, PI_I_e_rs, PI_I_e_nr, PI_I_e_nf,
    PI_I_k_Fe_ST2PS, PI_I_k_Fe_PS2ST, PI_I_k_Fe_ST2NR, PI_I_k_Fe_NR2ST, PI_I_k_Fe_ST2NF, PI_I_k_Fe_NF2ST,
    PI_I_KfePS, PI_I_KfeNR, PI_I_KfeNF, PI_I_KsatNH4, PI_I_KsatNO3, PI_I_KsatPO4, PI_I_KSAFe,
    PI_I_CHmax, PI_I_qNH4max, PI_I_qNO3max, PI_I_qPmax, PI_I_qFemax,
    PI_I_Chl2N, PI_I_R_NC, PI_I_R_PC, PI_I_NF_clock,
    PI_I_grz_P, PI_I_dvid_type, PI_I_dvid_P, PI_I_dvid_reg, PI_I_dvid_reg2, PI_I_mort_P, PI_I_mort_reg,
    PI_I_grazFracC, PI_I_grazFracN, PI_I_grazFracP, PI_I_grazFracFe,
    PI_I_mortFracC, PI_I_mortFracN, PI_I_mortFracP, PI_I_mortFracFe, PI_I_max_ptc,
    PI_NIPARAMS = 64
};

/* ---- abiotic particle parameters (Abiotic/particle_generation.jl:8,
 *      Parameters/param_default.jl:329-343) ----------------------------------- */
enum { PI_A_Nsuper = 0, PI_A_Rd, PI_A_release_P, PI_A_sz_min, PI_A_Ktr, PI_NAPARAMS = 5 };
#define PI_MAXSA 4          /* abiotic species per model */

/* ---- biogeochemical parameters (Parameters/param_default.jl:45-71) ------- */
enum {
    PI_B_kw = 0, PI_B_kc, PI_B_kDOC, PI_B_Nit, PI_B_kDON, PI_B_kDOP, PI_B_kPOC,
    PI_B_kPON, PI_B_kPOP, PI_B_kappa_h, PI_B_kappa_v, PI_B_kappa_hP, PI_B_kappa_vP,
    PI_B_shared_graz, PI_B_w_sink_org, PI_B_w_sink_inorg, PI_B_lambda_POC,
    PI_B_lambda_min, PI_B_kdiss, PI_B_lambda_dust, PI_B_lambda_Fe, PI_B_ligand,
    PI_B_DFeFrac,
    PI_NBGC = 23
};

/* ---- individual fields (QuotaMode StructArray columns, then PI_NEX extra
 *      float columns that only the modes with more state use) --------------- */
#define PI_NEX 8
enum {
    PI_F_x = 0, PI_F_y, PI_F_z, PI_F_Sz, PI_F_Bm, PI_F_Cq, PI_F_Nq, PI_F_Pq,
    PI_F_Chl, PI_F_gen, PI_F_age, PI_F_graz, PI_F_mort, PI_F_dvid,   /* float32 */
    PI_F_xi, PI_F_yi, PI_F_zi,                                       /* int64   */
    PI_F_ac,                                                         /* uint8   */
    PI_F_ex0,                                                        /* float32: PI_F_ex0 + k, k < PI_NEX */
    PI_NFIELDS = PI_F_ex0 + PI_NEX
};
/* CarbonMode (CarbonMode/plankton_generation.jl:1-12): the damaged biomass Bd
 * is the column PI_F_Cq; its Nq and Pq columns are zero. */
#define PI_F_Bd PI_F_Cq
/* MacroMolecularMode (MacroMolecularMode/plankton_generation.jl:1-12): protein,
 * DNA, carbohydrate/lipid and the N and P stores in the QuotaMode columns, RNA
 * in the first extra column. */
#define PI_F_PRO PI_F_Bm
#define PI_F_DNA PI_F_Sz
#define PI_F_CH  PI_F_Cq
#define PI_F_NST PI_F_Nq
#define PI_F_PST PI_F_Pq
#define PI_F_RNA PI_F_ex0
/* IronEnergyMode (IronEnergyMode/plankton_generation.jl:1-26): Sz, Bm, Chl as
 * in QuotaMode, the carbon, NH4 and P reserves in the Cq, Nq, Pq columns, the
 * rest in the extra columns; ptc (abiotic particles attached) stays 0 here. */
#define PI_F_qNH4  PI_F_Nq
#define PI_F_qP    PI_F_Pq
#define PI_F_qNO3  (PI_F_ex0 + 0)
#define PI_F_qFe   (PI_F_ex0 + 1)
#define PI_F_qFePS (PI_F_ex0 + 2)
#define PI_F_qFeNR (PI_F_ex0 + 3)
#define PI_F_qFeNF (PI_F_ex0 + 4)
#define PI_F_tdark (PI_F_ex0 + 5)
#define PI_F_exEn  (PI_F_ex0 + 6)
#define PI_F_ptc   (PI_F_ex0 + 7)

/* ---- grid-shaped outputs (timestepper + diagnostics) --------------------- */
enum {
    PI_G_par = 0,     /* timestepper.par   (padded) */
    PI_G_par0,        /* timestepper.par₀  (padded) */
    PI_G_Chl,         /* timestepper.Chl   (padded) accumulated Chl per cell */
    PI_G_pop,         /* timestepper.pop   (padded) individuals per cell */
    PI_G_temp,        /* timestepper.temp  (padded) */
    PI_G_PARF,        /* timestepper.PARF  (Nx,Ny), attenuated to the bottom after a step */
    PI_G_u, PI_G_v, PI_G_w,   /* timestepper.vel₁ (padded, halos filled) */
    PI_G_diag_PAR,    /* diags.tracer.PAR (padded, accumulated) */
    PI_G_plk_DIC      /* timestepper.plk.<tracer> : add the tracer id (padded) */
};
enum {                 /* per-species diagnostics (Diagnostics/diagnostics_struct.jl:97-104) */
    PI_D_num = 0, PI_D_graz, PI_D_mort, PI_D_dvid,
    PI_D_ptc,          /* particles attached, accumulated per step when the model has abiotic particles
                        * (time_step.jl:140-149; nonzero only in IronEnergyMode) */
    PI_NDIAG = 5
};

typedef struct pi_grid_desc {
    int32_t Nx, Ny, Nz;          /* interior size */
    int32_t Hx, Hy, Hz;          /* halo (reference default 2,2,2; >= 2 required by DST3FL) */
    int32_t topo_x, topo_y, topo_z; /* PI_PERIODIC / PI_BOUNDED; z must be bounded (tracer_sinking.jl) */
    float dx, dy;                /* RectilinearGrid Δx, Δy (rectilinear_grid.jl:72-73) */
    const float* dzC;            /* length Nz+2Hz, as built by rectilinear_grid.jl:93-116 */
    const float* dzF;            /* length Nz+2Hz */
    const uint8_t* landmask;     /* padded Bool landmask (Grids/utils.jl:92-129); NULL = no land */
} pi_grid_desc;

/* LatLonGrid (Grids/lat_lon_grid.jl:1-37, built by LatLonGrid(...) at :70-200
 * or LoadLatLonGrid(...) at :230-327): spacings, areas and volumes as arrays
 * with halos, column-major.  Replace Δx, Δy of pi_grid_desc in every formula
 * of Grids/areas_volumes_spacings.jl:24-38 (ΔxC(i,j,k) = dxC[i,j], Ax(i,j,k) =
 * Ax[i,j,k], volume(i,j,k) = Vol[i,j,k], ...); dzC, dzF stay in pi_grid_desc. */
typedef struct pi_metrics {
    const float *dxC, *dyC, *dxF, *dyF, *Az;   /* (Nx+2Hx, Ny+2Hy) */
    const float *Ax, *Ay, *Vol;                /* (Nx+2Hx, Ny+2Hy, Nz+2Hz) */
} pi_metrics;

typedef struct pi_model_desc {
    pi_grid_desc grid;
    int32_t n_species;
    int64_t capacity;            /* max_individuals per species */
    const float* params;         /* n_species * PI_NPARAMS floats */
    const float* bgc;            /* PI_NBGC floats */
    uint64_t rng_seed;           /* key of the shared uniform stream (pi_math.h) */
    int32_t n_diag_tracers;      /* tracers accumulated into diags.tracer besides PAR */
    const int32_t* diag_tracers; /* tracer ids, may be NULL when n_diag_tracers == 0 */
    int32_t mode;                /* PI_MODE_QUOTA, _CARBON, _MACROMOLECULAR or _IRONENERGY (PlanktonModel(...; mode)) */
    const pi_metrics* metrics;   /* LatLonGrid metrics; NULL = RectilinearGrid (dx, dy of grid) */
    /* abiotic particles (abiotic_setup, src/PlanktonIndividuals.jl:124-129;
     * Individuals/Abiotic/particle_generation.jl:1-35): n_abiotic species of
     * `capacity` slots each, PI_NAPARAMS floats per species.  They are carried
     * by the flow and the random walk, created at the faces by their flux
     * boundary conditions and counted per cell. */
    int32_t n_abiotic;
    const float* abiotic_params;
    /* phytoplankton-particle interaction (Abiotic/particle_interaction.jl:1-145,
     * time_step.jl:120-150): the pairs of palat.intac and palat.release as
     * (species, abiotic species) int32 pairs, 0-based, at most 4 of each, run
     * in this order every step after the physiology.  Interaction needs
     * IronEnergyMode (its max_ptc and ptc column); release works in every mode
     * (ptc of the other modes stays 0).  max_candidates (models.jl:60, 25 by
     * default) in [1, 64]. */
    int32_t n_interactions;
    const int32_t* interactions;
    int32_t n_releases;
    const int32_t* releases;
    int32_t max_candidates;
} pi_model_desc;

/* Host-side view of one species' individuals, all arrays of length capacity.
 * Pointers may be NULL on upload for fields that are zero. */
typedef struct pi_individuals_host {
    float *x, *y, *z, *Sz, *Bm, *Cq, *Nq, *Pq, *Chl, *gen, *age, *graz, *mort, *dvid;
    int64_t *xi, *yi, *zi;
    uint8_t *ac;
    float *ex[PI_NEX];           /* the extra columns (PI_F_ex0 + k); NULL when the mode has none */
} pi_individuals_host;

typedef struct pi_model pi_model;

/* Library / device set-up.  Replaces `isfunctional(arch)` (Architectures.jl:24). */
pi_status pi_init(int32_t device);
const char* pi_last_error(void);
const char* pi_version(void);

/* PlanktonModel(arch, grid; ...) storage (Model/models.jl:51-155, timestepper.jl:23-60).
 * The model starts at iteration 0 with zero tracers/velocities; upload the
 * initial state with the calls below. */
pi_status pi_model_create(const pi_model_desc* desc, pi_model** out);
pi_status pi_model_destroy(pi_model* m);

/* Initial state (generate_individuals / generate_tracers, Individuals.jl:42-72,
 * tracer_fields.jl:93-134): full-capacity SoA for species `sp` (the mode's
 * columns: MacroMolecularMode fills PI_F_PRO ... and ex[0] = RNA). */
pi_status pi_upload_individuals(pi_model* m, int32_t sp, const pi_individuals_host* h);
/* The same initial state generated on the device (no host arrays): every
 * species' slots as initialize_plankton! (QuotaMode/plankton_generation.jl:35-72,
 * CarbonMode/plankton_generation.jl:34-56, MacroMolecularMode/plankton_generation.jl:31-78)
 * with n_individuals[sp] active, and the tracers as generate_tracers with the
 * (initial_condition, rand_noise) source (tracer_fields.jl:93-134): value
 * c0 * (lower + k * 1f-4), k < n drawn per padded cell, then fill_halo_tracer!. */
pi_status pi_generate_state(pi_model* m, const int64_t* n_individuals, const float* tracer_c0,
                            const float* tracer_lower, const int32_t* tracer_n);
/* One padded tracer field (with halos). */
pi_status pi_upload_tracer(pi_model* m, int32_t tracer, const float* padded);

/* Forcing inputs staged by update! (Simulation/update.jl:18-26):
 *   velocity interior slices like vel_copy! (Simulation/utils.jl:5-11):
 *     u: (Nx[+1 if bounded x], Ny, Nz), v: (Nx, Ny[+1], Nz), w: (Nx, Ny, Nz[+1])
 *   PARF: (Nx, Ny)         temperature interior: (Nx, Ny, Nz)
 * A call replaces vel₁ / PARF / temp for the following steps. */
pi_status pi_set_velocity(pi_model* m, const float* u, const float* v, const float* w);
pi_status pi_set_parf(pi_model* m, const float* parf);
pi_status pi_set_temperature(pi_model* m, const float* temp);

/* TimeStep!(model, ΔT, diags) (Model/time_step.jl:12-180).  Enqueued on the
 * model's stream; an error raised on the device (capacity) is reported by this
 * call or, at the latest, by the next pi_sync / download. */
pi_status pi_step(pi_model* m, float dt);
/* Run `n` steps (update! without staging changes). */
pi_status pi_run(pi_model* m, float dt, int64_t n);
pi_status pi_sync(pi_model* m);

/* Outputs (Output/write_outputs.jl:159-202 read these).  pi_download_field
 * refuses an extra column (PI_F_ex0 + k) that the model's mode does not have. */
pi_status pi_download_field(pi_model* m, int32_t sp, int32_t field, void* host, int64_t n);
pi_status pi_download_tracer(pi_model* m, int32_t tracer, float* padded);
pi_status pi_download_grid(pi_model* m, int32_t which, float* host, int64_t n);
pi_status pi_download_diag(pi_model* m, int32_t sp, int32_t which, float* padded);
pi_status pi_download_diag_tracer(pi_model* m, int32_t tracer, float* padded);
/* Zero every diagnostic accumulator (write_diags_to_jld2, write_outputs.jl:186-201). */
pi_status pi_reset_diags(pi_model* m);
/* Choose the tracers accumulated into diags.tracer besides PAR
 * (PlanktonDiagnostics(model; tracer=...), Diagnostics/diagnostics_struct.jl:55-84);
 * resets those accumulators. */
pi_status pi_set_diag_tracers(pi_model* m, int32_t n, const int32_t* ids);

/* Flux boundary conditions of the tracers: set_bc!(model; tracer, pos, bc_value)
 * (HaloRegions/boundary_conditions.jl:12-29), applied to the tracer tendency
 * by apply_bcs! (HaloRegions/apply_bcs.jl:1-99) after the forcing and before
 * apply_tendency!: per tracer, in the order west, east, south, north, bottom,
 * top, Gc += bc·ΔT·A/V at the first interior cell of west/south/bottom and
 * Gc −= bc·ΔT·A/V at the last of east/north/top (A = Ax, Ay, Az of that cell,
 * V its volume).  `kind` follows getbc's dispatch (boundary_conditions.jl:
 * 31-40): PI_BCK_NONE removes it (`nothing`), PI_BCK_CONST takes values[0] (a
 * Number), PI_BCK_FIELD a face array (a 2-D array: west/east (Ny,Nz),
 * south/north (Nx,Nz), bottom/top (Nx,Ny), column-major), PI_BCK_SERIES nt
 * face arrays one after the other (a 3-D array (face..., nΔT) read at the
 * model's iteration: a step whose iteration exceeds nt is refused with
 * PI_ERR_ARG, where the reference reads out of bounds).  Values are copied. */
enum { PI_BC_WEST = 0, PI_BC_EAST, PI_BC_SOUTH, PI_BC_NORTH, PI_BC_BOTTOM, PI_BC_TOP, PI_NBC };
enum { PI_BCK_NONE = 0, PI_BCK_CONST, PI_BCK_FIELD, PI_BCK_SERIES };
pi_status pi_set_tracer_bc(pi_model* m, int32_t tracer, int32_t face, int32_t kind, const float* values,
                           int64_t nt);

/* Abiotic particles (Abiotic/particle_generation.jl): the columns of species
 * `sa` (x, y, z: float, xi, yi, zi: int64, ac: uint8; merg and idx are always
 * 0 here), each of length capacity.  Upload: x, y, z, ac (initialize_abiotic_particle!,
 * :24-35, drawn on the host); xi, yi, zi start at 0. */
enum { PI_AB_x = 0, PI_AB_y, PI_AB_z, PI_AB_xi, PI_AB_yi, PI_AB_zi, PI_AB_ac };
pi_status pi_upload_abiotic(pi_model* m, int32_t sa, const float* x, const float* y, const float* z,
                            const uint8_t* ac);
pi_status pi_download_abiotic(pi_model* m, int32_t sa, int32_t field, void* host, int64_t n);
/* diags.abiotics.saN.num (padded): particles per cell, accumulated per step
 * (time_step.jl:152-159); zeroed by pi_reset_diags. */
pi_status pi_download_abiotic_diag(pi_model* m, int32_t sa, float* padded);
/* set_bc_particle!(model; sa, pos, bc_value) (boundary_conditions.jl:31-39):
 * a flux of particles through a face (same positions, kinds and nt rule as
 * pi_set_tracer_bc); particles_from_bcs! (particle_interaction.jl:249-267)
 * creates one particle at the cell next to the face when a uniform draw times
 * sz_min·Nsuper is below bc·ΔT·A. */
pi_status pi_set_particle_bc(pi_model* m, int32_t sa, int32_t face, int32_t kind, const float* values, int64_t nt);

/* Scalars. */
pi_status pi_get_iteration(pi_model* m, int64_t* iteration);
pi_status pi_count_active(pi_model* m, int32_t sp, int64_t* n);   /* dot(ac, ac) (models.jl:157-164) */
/* Σ over the steps taken so far of the individuals each step moved and grew
 * (all species, alive at the step's gather): the numerator of
 * individual-steps/s.  No counterpart in the reference (its benchmarks time
 * whole steps).  Synchronises the model's stream. */
pi_status pi_individual_steps(pi_model* m, int64_t* n);

/* Per-kernel timing with HIP events recorded on the model's stream around each
 * launch group; pi_kernel_times returns, per group name (static text), the total
 * milliseconds and the number of launches since timing was enabled or last
 * read.  No counterpart in the reference (its benchmarks time whole steps). */
pi_status pi_enable_timing(pi_model* m, int32_t on);
pi_status pi_kernel_times(pi_model* m, int32_t* n, const char** names, float* ms, int32_t* counts, int32_t cap);

/* Launch shapes: how the work of a step is split over workgroups.  No
 * counterpart in the reference (KernelAbstractions picks its own workgroup
 * sizes).  Every accepted value gives the same results bit for bit; only the
 * speed changes, so a search by timing steps (the package's autotune) may set
 * them per model.  A model in a batch keeps its own values: run lengths apply
 * per member; block counts are those of its share of a batched launch. */
enum {
    PI_S_XY_RUN = 0,       /* rows of one level per work item of the fused x/y sweep (1..1024; 16) */
    PI_S_FINAL_RUN,        /* levels per column run of the z sweep + update (1..4096; 16) */
    PI_S_ADV_BLOCKS,       /* workgroups of the advection launch, at most (1..2^24; 16384) */
    PI_S_PHYS_BLOCKS,      /* workgroups of the physiology launch, at most (1..2^24; 4096) */
    PI_S_CELLSUM_PER_CU,   /* workgroups per compute unit of the per-cell Chl/pop sums (1..64; 8) */
    PI_S_PLK_BLOCKS,       /* workgroups of the per-cell exchange sums, at most (1..2^24; 2^22) */
    PI_NSHAPE
};
pi_status pi_set_launch_shape(pi_model* m, int32_t which, int64_t value);
pi_status pi_get_launch_shape(const pi_model* m, int32_t which, int64_t* value);

/* Device memory.  pi_footprint: the bytes a model with this descriptor would
 * allocate (a dry run of pi_model_create that touches no device);
 * pi_model_footprint: the bytes a model holds.  No counterpart in the
 * reference (Julia allocates per array); used to admit ensemble members by
 * footprint. */
pi_status pi_footprint(const pi_model_desc* desc, int64_t* bytes);
pi_status pi_model_footprint(const pi_model* m, int64_t* bytes);

/* Batches: many models stepped by the same launches.  No counterpart in the
 * reference, which steps one PlanktonModel per update!/TimeStep! call
 * (Simulation/update.jl:28, Model/time_step.jl:12); a batch steps each member
 * exactly as pi_run would, all members together: at every kernel of the step
 * one launch covers the blocks of all members.  A member keeps its buffers and
 * its API (staging, downloads) while in a batch; its calls then go to the
 * batch's stream.  pi_batch_run(b, dt, n) = n steps of every member (members
 * in an error state are skipped); its errors are reported per member by
 * pi_sync / the downloads.  Members may be added and removed between runs. */
typedef struct pi_batch pi_batch;
pi_status pi_batch_create(pi_batch** out);
pi_status pi_batch_add(pi_batch* b, pi_model* m);
pi_status pi_batch_remove(pi_batch* b, pi_model* m);
pi_status pi_batch_run(pi_batch* b, float dt, int64_t n);
/* Blocks per batched launch of a block-stride kernel, shared among the members
 * in proportion to their own grids (default 16384).  Any value gives the same
 * results; Ensemble.autotune picks it by timing steps. */
pi_status pi_batch_set_budget(pi_batch* b, int64_t blocks);
/* pi_batch_run that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under stream capture,
 * serving loops that forbid hipMalloc).  Same steps and results as
 * pi_batch_run, every step launched eagerly (no graph), with the batched
 * launches' argument and block tables in one static device array of 32 MiB
 * per GPU instead of device allocations.  A batch whose tables exceed it is
 * refused with PI_ERR_NOMEM before anything runs.  Calls on one GPU share
 * that array: calls that overlap in time on one GPU race. */
pi_status pi_batch_run_noalloc(pi_batch* b, float dt, int64_t n);
/* Steady steps replay a captured graph unless this is off: turn it off when
 * other threads make HIP calls (model creation, downloads) while the batch
 * runs, since a capture forbids them. */
pi_status pi_batch_graphs(pi_batch* b, int32_t on);
pi_status pi_batch_sync(pi_batch* b);
pi_status pi_batch_size(const pi_batch* b, int32_t* n);
pi_status pi_batch_destroy(pi_batch* b);

/* Measurement aid, no counterpart in the reference: copies `bytes` (a multiple
 * of 16) between two fresh device buffers with the kernel k_probe_copy on the
 * current device and waits for it.  Hardware counters of that dispatch scale
 * the HBM traffic counters of the other kernels (bytes read = bytes written =
 * `bytes`). */
pi_status pi_probe_copy(int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* PI_HIP_H */
