/*
 * CPU oracle of the TNL-LBM lattice update -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker: the product path (the HIP
 * library libtnl_lbm_amd.so) never calls into it.
 *
 * It is a restatement, operation order preserved, of the reference's host
 * instantiation (TNL-LBM include/lbm3d/kernels.h:59-196 looped as in
 * include/lbm3d/state.hpp:1388-1397).  Each function cites the reference
 * file:line it restates.  The entry points take the same POD structs as
 * the C ABI (include/tnl_lbm_amd.h) but with HOST pointers.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include "../include/tnl_lbm_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* oracle_last_error(void);

/* whole-lattice operations (host memory) */
int oracle_d3q27_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const int64_t offset[3],
                      const int64_t end[3], uint32_t flags, int nthreads);
int oracle_nse_ade_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* nse, const tnl_lbm_d3q7_data* ade,
                        const int64_t offset[3], const int64_t end[3], uint32_t flags, int nthreads);
int oracle_d3q27_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, double rho, double vx,
                                 double vy, double vz);
int oracle_d3q27_set_equilibrium_field(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const double* rho,
                                       const double* vx, const double* vy, const double* vz);
int oracle_d3q27_compute_macro(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD);
int oracle_d3q27_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which, double* out);
int oracle_d3q27_import_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, int which, const double* in);
int oracle_d3q27_adjoint_step(const tnl_lbm_config* cfg, const tnl_lbm_d3q27_data* SD, const void* lambda_next,
                              void* lambda, double* grad_inflow);
int oracle_d3q7_set_equilibrium(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE, const void* phi,
                                double phi0);
int oracle_d3q7_export_f64(const tnl_lbm_config* cfg, const tnl_lbm_d3q7_data* ADE, int which, double* out);

/* single-site operations, all in double (or float when `single` != 0)   */
int oracle_d3q27_collide(int collision, int single, double* f, double nu, double fx, double fy, double fz,
                         double* rho_u_out);
int oracle_d3q27_collide_dev(int collision, double* g, double nu, double fx, double fy, double fz, int single);
int oracle_d3q27_equilibrium(int which_eq, int single, double rho, double vx, double vy, double vz, double* f);
int oracle_d3q27_density_velocity(int single, const double* f, double fx, double fy, double fz, double* out4);

/* 16-bit conversions used by the deviation storage forms */
uint16_t oracle_f32_to_f16(float v);
float oracle_f16_to_f32(uint16_t h);
uint16_t oracle_f32_to_bf16(float v);
float oracle_bf16_to_f32(uint16_t h);

#ifdef __cplusplus
}
#endif
#endif
