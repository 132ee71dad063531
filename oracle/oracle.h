/*
 * oracle.h -- CPU restatement of Sedona's implicit Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This plain-C FP64 code restates, function by
 * function, the reference algorithms that the HIP library replaces; it is the
 * checker the parity tests compare the GPU path against and the CPU baseline
 * that bench.py times.  Nothing in the product links, loads or calls it.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the Sedona source tree).  The data descriptors are the C-ABI structs of
 * include/sedona_gpu.h, here holding HOST pointers.
 *
 * Parity pins (see tests/test_oracle_*.py): every function is checked against
 * closed forms, invariants, brute force on tiny inputs or the reference's own
 * fixtures; the functions for which no such check exists are listed as
 * "parity unpinned" in the repository's design notes.
 */
#ifndef SEDONA_ORACLE_H
#define SEDONA_ORACLE_H

#include <stdint.h>
#include "../include/sedona_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* physical constants, verbatim values of src/utils/physical_constants.h:9-33
 * (including the truncated pi = 3.14159 and alpha_fs = 0.007) */
#define OR_PI          3.14159
#define OR_C           2.99792458e10
#define OR_H           6.6260755e-27
#define OR_K           1.380658e-16
#define OR_K_EV        8.6173324e-5
#define OR_M_P         1.67262158e-24
#define OR_M_E         9.10938188e-28
#define OR_SB          5.6704e-5
#define OR_A_RAD       7.5657e-15
#define OR_M_E_MEV     0.510998910511
#define OR_THOMSON     0.66523e-24
#define OR_SIGMA_TOT   0.0265400193567
#define OR_ALPHA_FS    0.007
#define OR_EV_TO_ERGS  1.60217646e-12
#define OR_MEV_TO_ERGS 1.60217646e-6
#define OR_M_SUN       1.98892e33

/* radioactive constants, src/transport/radioactive.h:11-35 */
#define OR_TAU_56NI  757241.8
#define OR_TAU_56CO  9627379.
#define OR_E_56NI    1.725
#define OR_E_56CO    3.730
#define OR_CO_POS    0.0324
#define OR_TAU_48CR  111976.
#define OR_TAU_48V   1.99107843e6
#define OR_E_48CR    0.442
#define OR_E_48V     3.066

/* ---- counter-based RNG (Philox-4x32-10, Salmon et al. SC'11) ---- */
void     or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double   or_uniform(uint64_t seed, uint64_t id, uint32_t draw);
float    or_uniform_f32(uint64_t seed, uint64_t id, uint32_t draw);

/* ---- locate_array (src/utils/locate_array.cpp) ---- */
int      or_log_init(double start, double stop, double del, double* x, int max_n, double* min_out);
int      or_lin_init(double start, double stop, double del, double* x, int max_n, double* min_out);
int      or_locate(const sgpu_locate* g, double xval);
int      or_locate_within_bounds(const sgpu_locate* g, double xval);
double   or_center(const sgpu_locate* g, int i);
double   or_delta(const sgpu_locate* g, int i);
double   or_left(const sgpu_locate* g, int i);
double   or_sample(const sgpu_locate* g, int i, double r);
int      or_cdf_sample(const double* y, int n, double yval);
void     or_cdf_normalize(double* y, int n);

/* ---- gas state and opacity (src/opacity, src/transport/transport_opacity.cpp) ---- */
void     or_decay_composition(int n_el, const int32_t* Z, const int32_t* A, double* X, double t);
double   or_decay(int n_el, const int32_t* Z, const int32_t* A, const double* X, double t,
                  double* gfrac, int force_rproc);
void     or_solve_lte(const sgpu_atoms* at, int a, double ne, double T,
                      double* ion_part, double* ion_frac, double* lev_n);
int      or_solve_state(const sgpu_atoms* at, int n_el, const int32_t* elem_A,
                        const double* mass_frac, double mu_I, double dens, double T,
                        double* n_elec, double* ion_part, double* ion_frac, double* lev_n,
                        int* n_evals);
int      or_opacity_build(const sgpu_locate* nu_grid, const sgpu_zone_state* zs,
                          const sgpu_atoms* at, const sgpu_opacity_params* p,
                          sgpu_opacity_out* out);
int      or_opacity_build_nlte(const sgpu_locate* nu, const sgpu_zone_state* zs, const sgpu_atoms* at,
                               const sgpu_opacity_params* p, const double* J_nu, const int32_t* atom_nlte,
                               int use_collisions, sgpu_opacity_out* out);
double   or_klein_nishina(double x);
double   or_blackbody_nu(double T, double nu);
void     or_solve_eq_temperature(const sgpu_locate* nu, int nz, const double* abs_opacity,
                                 const double* e_abs, const double* e_rad, double tmin, double tmax,
                                 int set_Tgas_to_Trad, double* T_gas, int* n_eval);

/* ---- transport (src/transport, src/grid) ---- */
int      or_get_zone(const sgpu_grid* g, const double* x);
int      or_get_next_zone(const sgpu_grid* g, const double* x, const double* D, int i,
                          double r_core, double* l);
int      or_emit_radioactive(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                             const sgpu_zone_state* zs, const sgpu_tables* tab,
                             const sgpu_transport_params* p, int64_t n_emit,
                             int force_rproc, uint64_t id_base,
                             sgpu_packets* pk, int64_t offset, sgpu_tallies* ta,
                             int64_t* n_out);
int      or_initialize_particles(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                                 const sgpu_zone_state* zs, const double* e_rad, double* emis_cdf,
                                 const sgpu_transport_params* p, int64_t n_init, uint64_t id_base,
                                 sgpu_packets* pk, int64_t offset, int64_t* n_out);
int      or_emit_inner_source(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                              const sgpu_core_source* core, const sgpu_transport_params* p,
                              int64_t n_emit, uint64_t id_base, sgpu_packets* pk, int64_t offset,
                              int64_t* n_out);
int      or_propagate(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                      const sgpu_tables* tab, const sgpu_transport_params* p,
                      sgpu_packets* pk, int64_t n, sgpu_tallies* ta,
                      sgpu_counters* cnt, int n_threads);
int      or_normalize_tallies(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                              const sgpu_transport_params* p, sgpu_tallies* ta);
int64_t  or_compact(sgpu_packets* pk, int64_t n, int64_t* n_escaped);
void     or_homologous_step(sgpu_grid* grid, double* rho, double* r_out, double* x_out,
                            double* z_out, double* vol, double e);
/* test hooks on single packets */
void     or_lorentz(const sgpu_grid* grid, double* x, double* D, double* e, double* nu, int tolab);
void     or_compton_sample(double nu, int64_t n, uint64_t seed, double* cost_out, double* eratio_out);

/* ---- NLTE (src/opacity/AtomicSpecies_state.cpp:131-196) ---- */
int      or_nlte_solve(int n, const double* rates, const double* lte, double* b, double* pop);
int      or_nlte_set_rates(const sgpu_atoms* at, int a, double ne, double T, const double* line_J,
                           const double* Pic, const double* Rci, const double* lev_lte,
                           int use_collisions, int no_ground_recomb, double* rates, double* lte);
void     or_nlte_populations(const sgpu_atoms* at, int a, const double* b, const double* lte,
                             double* lev_n, double* ion_frac);
double   or_voigt_profile(double x, double a);
int      or_continuum_level(const sgpu_atoms* at, int a, int k);
double   or_photo_cs(double E_ion, double chi_ion, double E_lev, int ion, double E_ev);
void     or_nlte_radiative_rates(const sgpu_atoms* at, int a, const sgpu_locate* nu, const double* J_nu,
                                 double T, int elem_A, double line_velocity_width, double* Pic,
                                 double* Rci, double* line_J);
int      or_lu_decomp(int n, double* A, int* perm);
void     or_lu_solve(int n, const double* LU, const int* perm, const double* rhs, double* x);

/* ---- differentiable transport (continuous absorption) ---- */
int      or_adjoint_forward(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                            const sgpu_locate* spec_grid, const float* kappa_abs,
                            const double* rho, const double* scat,
                            const sgpu_transport_params* p, const sgpu_packets* pk,
                            int64_t n, double* F, int32_t* rec_bin, double* rec_w,
                            double* dF_dkappa_dir, const double* dir);
int      or_adjoint_backward(const sgpu_grid* grid, const sgpu_locate* nu_grid,
                             const sgpu_locate* spec_grid, const float* kappa_abs,
                             const double* rho, const double* scat,
                             const sgpu_transport_params* p, const sgpu_packets* pk,
                             int64_t n, const double* dL_dF, const int32_t* rec_bin,
                             const double* rec_w, double* grad);

#ifdef __cplusplus
}
#endif
#endif
