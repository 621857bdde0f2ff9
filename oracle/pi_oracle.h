/*
 * pi_oracle.h - TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference PlanktonIndividuals.jl CPU() time step
 * (QuotaMode and CarbonMode, tracer flux boundary conditions, abiotic particles without interaction), used as the checker for libpi_hip.so.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product never does.
 *
 * Same descriptor types and field ids as include/pi_hip.h, prefix orc_.
 * precision: 32 = FP32 in the reference's operation order (bit-parity mode),
 *            64 = the same code in binary64 with Float32-rounded constants
 *                 (error reference).
 */
#ifndef PI_ORACLE_H
#define PI_ORACLE_H
#include "../include/pi_hip.h"

#ifdef __cplusplus
extern "C" {
#endif
typedef struct orc_model orc_model;

const char* orc_last_error(void);
pi_status orc_model_create(const pi_model_desc* desc, int32_t precision, orc_model** out);
pi_status orc_model_destroy(orc_model* m);
pi_status orc_upload_individuals(orc_model* m, int32_t sp, const pi_individuals_host* h);
pi_status orc_upload_tracer(orc_model* m, int32_t tracer, const float* padded);
pi_status orc_set_velocity(orc_model* m, const float* u, const float* v, const float* w);
pi_status orc_set_parf(orc_model* m, const float* parf);
pi_status orc_set_temperature(orc_model* m, const float* temp);
pi_status orc_step(orc_model* m, float dt);
pi_status orc_set_tracer_bc(orc_model* m, int32_t tracer, int32_t face, int32_t kind, const float* values,
                            int64_t nt);
pi_status orc_set_particle_bc(orc_model* m, int32_t sa, int32_t face, int32_t kind, const float* values, int64_t nt);
pi_status orc_upload_abiotic(orc_model* m, int32_t sa, const float* x, const float* y, const float* z,
                             const uint8_t* ac);
pi_status orc_download_abiotic(orc_model* m, int32_t sa, int32_t field, void* host, int64_t n);
pi_status orc_download_abiotic_diag(orc_model* m, int32_t sa, float* padded);
/* Downloads convert to float32 in FP64 mode; *_f64 variants give binary64. */
pi_status orc_download_field(orc_model* m, int32_t sp, int32_t field, void* host, int64_t n);
pi_status orc_download_field_f64(orc_model* m, int32_t sp, int32_t field, double* host, int64_t n);
pi_status orc_download_tracer(orc_model* m, int32_t tracer, float* padded);
pi_status orc_download_tracer_f64(orc_model* m, int32_t tracer, double* padded);
pi_status orc_download_grid(orc_model* m, int32_t which, float* host, int64_t n);
pi_status orc_download_diag(orc_model* m, int32_t sp, int32_t which, float* padded);
pi_status orc_download_diag_tracer(orc_model* m, int32_t tracer, float* padded);
pi_status orc_reset_diags(orc_model* m);
pi_status orc_set_diag_tracers(orc_model* m, int32_t n, const int32_t* ids);
pi_status orc_get_iteration(orc_model* m, int64_t* iteration);
pi_status orc_count_active(orc_model* m, int32_t sp, int64_t* n);
pi_status orc_individual_steps(orc_model* m, int64_t* n);

/* Building blocks exposed for the oracle's own pin tests. */
pi_status orc_fill_halo(int32_t Nx, int32_t Ny, int32_t Nz, int32_t Hx, int32_t Hy, int32_t Hz,
                        int32_t tx, int32_t ty, int32_t tz, float* c);
/* one halo statement of halo_regions.jl on a padded float field: op 0 = fill_halo_west!/
 * south!/top! (dim 0/1/2), 1 = east!/north!/bottom!, 2 = *_vel!, 3 = *_Gc!,
 * 4 = fill_halo_top_flux_sink!, 5 = fill_halo_bottom_flux_sink!; topo = PI_PERIODIC/PI_BOUNDED */
pi_status orc_fill_halo_op(int32_t op, int32_t dim, int32_t topo, int32_t Nx, int32_t Ny, int32_t Nz,
                           int32_t Hx, int32_t Hy, int32_t Hz, float* c);
/* threads of the element-wise loops (OpenMP; default 1); returns the value set */
int32_t orc_set_threads(int32_t n);
float orc_expf(float x);
double orc_exp_f64(double x);
void orc_log_f64_v(const double* x, double* y, int64_t n);
void orc_cos2pi_f64_v(const double* x, double* y, int64_t n);
void orc_init_normal_v(const float* a, const float* b, float* y, int64_t n);
void orc_uniform3(uint64_t seed, int64_t iteration, int32_t sp, int32_t kind, int64_t slot, float* out3);
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_tempfunc(const float* params, float T);
float orc_tempfunc_ps(const float* params, float T);
#ifdef __cplusplus
}
#endif
#endif
