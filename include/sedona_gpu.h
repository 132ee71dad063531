/*
 * sedona_gpu.h -- C-ABI of the MI355X implicit-Monte-Carlo transport hot path.
 *
 * Plain C: extern "C", plain pointers and sizes, no exceptions across the
 * boundary, no HIP or torch types in the signatures.  Every compute entry
 * point is stream-ordered on the context's stream and is called from one host
 * thread; it replaces a whole reference loop, never a per-packet call.
 *
 * Memory: every array pointer handed to an sgpu_* compute call is a DEVICE
 * pointer (allocate with sgpu_alloc, hipMalloc or a torch tensor).  Tables
 * are contiguous and zone-major: table[zone * n_nu + bin].
 *
 * Status codes follow the reference's conventions:
 *   GasState::solve_state returns 0 ok / 1 root not bracketed / 2 max iter
 *   (src/opacity/GasState.cpp:228-231, 281-282, 335); negative values are
 *   argument or HIP errors of this library; sgpu_last_error() gives the text.
 *
 * Reference interfaces replaced (paths relative to the Sedona source tree):
 *   sgpu_opacity_build      transport::set_opacity            src/transport/transport_opacity.cpp:21-260
 *                           (GasState::solve_state + computeOpacity +
 *                            get_planck_mean / get_rosseland_mean)
 *                                                             src/opacity/GasState.cpp:208-337
 *                                                             src/opacity/GasState_opacities.cpp:11-136,464-494,564-666
 *   sgpu_emit_radioactive   transport::emit_radioactive       src/transport/emission.cpp:165-243
 *   sgpu_initialize_particles transport::initialize_particles src/transport/emission.cpp:112-159
 *   sgpu_emit_inner_source  transport::emit_inner_source      src/transport/emission.cpp:370-480
 *                           (+ setup_core_emission            src/transport/transport_init.cpp:465-548)
 *   sgpu_propagate          the packet loop of transport::step src/transport/transport.cpp:57-88
 *                           + propagate / propagate_monte_carlo :202-484
 *   sgpu_compact            transport::clean_up_particle_vector src/transport/transport.cpp:164-194
 *   sgpu_compact_noalloc    (no reference counterpart: sgpu_compact for callers that cannot allocate)
 *   sgpu_wipe_radiation     transport::wipe_radiation          src/transport/transport_mpi.cpp:18-34
 *   sgpu_normalize_tallies  transport::reduce_radiation (normalisation part)
 *                                                             src/transport/transport_mpi.cpp:415-441
 *   sgpu_nlte_solve         AtomicSpecies::solve_nlte (matrix assembly + LU)
 *                                                             src/opacity/AtomicSpecies_state.cpp:131-196
 *   sgpu_nlte_batched       (no reference counterpart: all systems of a step in one launch)
 *   sgpu_homologous_step    hydro_homologous::step + grid expand src/hydro/hydro_homologous.cpp:27-61
 *   sgpu_solve_eq_temperature transport::solve_eq_temperature src/transport/solve_equilibrium.cpp:66-320
 *   sgpu_nlte_radiative_rates AtomicSpecies::calculate_radiative_rates src/opacity/AtomicSpecies_state.cpp:202-287
 *   sgpu_nlte_set_rates     AtomicSpecies::set_rates            src/opacity/AtomicSpecies_state.cpp:294-432
 *   sgpu_nlte_populations   end of AtomicSpecies::solve_nlte    src/opacity/AtomicSpecies_state.cpp:178-193
 *   sgpu_adjoint_*          (no reference counterpart: differentiable transport)
 */
#ifndef SEDONA_GPU_H
#define SEDONA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* status codes                                                        */
/* ------------------------------------------------------------------ */
#define SGPU_OK              0
#define SGPU_NOT_BRACKETED   1   /* GasState::ne_brent_method error 1 */
#define SGPU_MAX_ITER        2   /* GasState::ne_brent_method error 2 */
#define SGPU_ERR_ARG        -1
#define SGPU_ERR_HIP        -2
#define SGPU_ERR_SIZE       -3
#define SGPU_ERR_NOT_CONVERGED -4

/* packet type / fate, same numbering as src/transport/particle.h:9-10 */
#define SGPU_PHOTON   0
#define SGPU_GAMMARAY 1
#define SGPU_MOVING   0
#define SGPU_STOPPED  1
#define SGPU_ESCAPED  2
#define SGPU_ABSORBED 3

/* locate_array types (src/utils/locate_array.h:13) */
#define SGPU_LOC_FLEX 0
#define SGPU_LOC_LIN  1
#define SGPU_LOC_LOG  2
#define SGPU_LOC_NONE 3

/* grid kinds */
#define SGPU_GRID_1D_SPHERE 1
#define SGPU_GRID_2D_CYLN   2

/* propagation precision */
#define SGPU_PREC_FP64  0   /* plain drop-in: FP64 everywhere, one thread per
                               packet, per-event tally atomics             */
#define SGPU_PREC_MIXED 1   /* FP32 geometry/opacity, FP64 tallies        */
#define SGPU_PREC_FP64_BATCHED 2  /* the FP64 arithmetic of SGPU_PREC_FP64 (the
                               same packet paths), launched as the mixed form
                               is: persistent lanes refilling from the packet
                               list, zone tallies summed in registers over a
                               zone visit into LDS, J_nu per (zone, bin) run */

/* NLTE precision */
#define SGPU_NLTE_FP64        0   /* FP64 partial-pivot LU                 */
#define SGPU_NLTE_FP32_REFINE 1   /* FP32 LU + FP64 iterative refinement   */

/* ------------------------------------------------------------------ */
/* data descriptors (plain structs of device pointers and sizes)      */
/* ------------------------------------------------------------------ */

/* A binned axis, the data of locate_array (src/utils/locate_array.h:15-90):
 * x[i] is the RIGHT edge of bin i, min the left edge of bin 0; bins are
 * closed on the left, open on the right. */
typedef struct {
  int32_t n;          /* number of bins (= number of right edges)        */
  int32_t type;       /* SGPU_LOC_*                                      */
  double  min;        /* left edge of bin 0                              */
  double  del;        /* lin: width; log: fractional step (1+del ratio)  */
  const double* x;    /* [n] right edges                                 */
} sgpu_locate;

/* Homologous grid.  1D: r_out shells, zone i = shell i (grid_1D_sphere).
 * 2D: cylindrical (x = varpi, z), zone index ix*nz + iz (grid_2D_cyln). */
typedef struct {
  int32_t kind;        /* SGPU_GRID_*                                    */
  int32_t n_zones;
  double  t_now;       /* grid clock: v = x / t_now, dv/ds = 1/t_now      */
  double  r_core;      /* 1D inner absorbing core radius (0 = none)      */
  sgpu_locate r_out;   /* 1D shell outer radii                            */
  sgpu_locate x_out;   /* 2D varpi edges (nx bins)                        */
  sgpu_locate z_out;   /* 2D z edges (nz bins)                            */
  const double* vol;   /* [n_zones] zone volumes                          */
} sgpu_grid;

/* Per-zone gas state, SoA (zone record src/grid/zone.h:9-50). */
typedef struct {
  int32_t n_zones;
  int32_t n_elems;
  const double* rho;          /* [nz] g/cm^3                                      */
  const double* T_gas;        /* [nz] K                                           */
  const double* X;            /* [nz][n_elems] initial (undecayed) mass fractions */
  const double* e_gas;        /* [nz] or NULL: only read when fleck_alpha != 0    */
  const double* grey_opacity; /* [nz] total_grey_opacity (cm^2/g); 0 = non-grey   */
  const int32_t* elem_Z;      /* [n_elems]                                        */
  const int32_t* elem_A;      /* [n_elems]                                        */
} sgpu_zone_state;

/* Atomic data for the gas: one atom per element (GasState::atoms), levels
 * and ions flattened over atoms; the fully ionised stage that the reference
 * appends (src/opacity/AtomicData.cpp:429-446) is included by the caller as
 * the last ion / level of each atom.  Fuzz lines (fuzz_line_structure,
 * src/opacity/AtomicData.h:12-20) sorted by (atom, bin) with a CSR index. */
typedef struct {
  int32_t n_atoms;                 /* == n_elems of the zone state        */
  int32_t max_levels;              /* max levels of any atom              */
  int32_t n_levels;                /* levels over all atoms (lev_off[n_atoms]) */
  int32_t n_ions;                  /* ions over all atoms (ion_off[n_atoms])   */
  const int32_t* lev_off;          /* [n_atoms+1]                         */
  const int32_t* ion_off;          /* [n_atoms+1]                         */
  const double*  lev_E;            /* [n_lev] excitation energy (eV)      */
  const int32_t* lev_g;            /* [n_lev] statistical weight          */
  const int32_t* lev_ion;          /* [n_lev] ion stage within its atom   */
  const double*  ion_chi;          /* [n_ion] ionisation energy (eV)      */
  const int32_t* atom_zero_eps;    /* [n_atoms] 1: epsilon = 0 for atom   */
  int64_t n_lines;                 /* fuzz lines                          */
  const double*  line_nu;          /* [n_lines] Hz                        */
  const double*  line_gf;          /* [n_lines]                           */
  const double*  line_El;          /* [n_lines] lower level energy (eV)   */
  const int32_t* line_ion;         /* [n_lines] ion stage within atom     */
  const int32_t* line_bin;         /* [n_lines] nu-grid bin               */
  const int64_t* line_bin_off;     /* [n_atoms*(n_nu+1)] CSR: lines of (atom a, bin b)
                                      are [off[a*(n_nu+1)+b], off[a*(n_nu+1)+b+1]) */
  /* level-resolved (bound-bound) lines, AtomicLine (AtomicData.h:29-41) with
   * the derived quantities of AtomicData.cpp:349-375: nu = dE/h, B_ul, B_lu,
   * f_lu, bin = locate_within_bounds(nu).  Sorted by (atom, bin), keeping
   * each atom's line order within a bin; NULL / 0 when absent.            */
  int64_t n_blines;
  const int32_t* bline_l;          /* [n_blines] lower level (index over all atoms) */
  const int32_t* bline_u;          /* [n_blines] upper level                */
  const double*  bline_nu;         /* [n_blines] Hz                         */
  const double*  bline_f;          /* [n_blines] oscillator strength f_lu   */
  const double*  bline_A;          /* [n_blines] Einstein A_ul (s^-1)       */
  const double*  bline_Bul;        /* [n_blines] Einstein B_ul              */
  const double*  bline_Blu;        /* [n_blines] Einstein B_lu              */
  const int32_t* bline_bin;        /* [n_blines] nu-grid bin                */
  const int64_t* bline_bin_off;    /* [n_atoms*(n_nu+1)] CSR over (atom, bin) */
} sgpu_atoms;

/* Opacity-build parameters (names of sedona_defaults.lua / transport members). */
typedef struct {
  double t_now;                    /* transport::t_now_ (decay and expansion time) */
  double dt;                       /* step length (Fleck factor)                   */
  double epsilon;                  /* opacity_epsilon                              */
  double temp_min;                 /* limits_temp_min                              */
  double temp_max;                 /* limits_temp_max                              */
  double maximum_opacity;          /* opacity_maximum_opacity                      */
  double line_velocity_width;      /* line_velocity_width                          */
  double fleck_alpha;              /* transport_fleck_alpha                        */
  int32_t use_electron_scattering; /* opacity_electron_scattering                  */
  int32_t use_fuzz_expansion;      /* opacity_fuzz_expansion                       */
  int32_t omit_scattering;         /* opacity_no_scattering                        */
  int32_t omit_composition_decay;  /* dont_decay_composition                       */
  int32_t radiative_eq;            /* transport_radiative_equilibrium              */
  int32_t first_step;              /* transport::first_step_: the state solve and
                                      opacities then use the UNclamped T_gas
                                      (transport_opacity.cpp:109-120)            */
  int32_t use_line_expansion;      /* opacity_line_expansion: level-resolved lines
                                      (GasState::line_expansion_opacity)          */
  int32_t use_free_free;           /* opacity_free_free (GasState::free_free_opacity) */
  int32_t use_bound_free;          /* opacity_bound_free (GasState::bound_free_opacity,
                                      hydrogenic cross-sections)                  */
  int32_t no_ground_recomb;        /* opacity_no_ground_recomb                     */
  int32_t use_bound_bound;         /* opacity_bound_bound: Voigt-profile line opacity
                                      (GasState_opacities.cpp:83-91,379-408,
                                      AtomicSpecies_opacities.cpp:224-282)      */
  double minimum_extinction;       /* opacity_minimum_extinction: bound-bound lines
                                      with alpha_0 / (nu_0^2 dnu) below it are
                                      skipped (AtomicSpecies_opacities.cpp:259,
                                      set at transport_init.cpp:225-226)        */
} sgpu_opacity_params;

/* Opacity-build outputs.  Tables [nz][n_nu]; scalars [nz]; optional
 * per-(zone, ion) / per-(zone, level) state arrays may be NULL. */
typedef struct {
  double* abs;              /* abs_opacity_ (cm^-1)                              */
  double* scat;             /* scat_opacity_                                     */
  double* emis_cdf;         /* emissivity_ CDF, normalised to 1                  */
  double* L_thermal;        /* zone L_thermal                                    */
  double* planck_mean;      /* planck_mean_opacity_                              */
  double* rosseland_mean;   /* rosseland_mean_opacity_                           */
  double* n_elec;           /* zone n_elec                                       */
  double* compton_opac;     /* gamma-ray Compton coefficient                     */
  double* photoion_opac;    /* gamma-ray photo-absorption coefficient            */
  double* eps_imc;          /* Fleck factor                                      */
  int32_t* zone_status;     /* solve_state return value per zone                 */
  double* ion_frac;         /* [nz][n_ion] or NULL                               */
  double* ion_part;         /* [nz][n_ion] or NULL                               */
  double* lev_n;            /* [nz][n_lev] or NULL                               */
} sgpu_opacity_out;

/* Monte Carlo packets, SoA (class particle, src/transport/particle.h:12-48).
 * id and draw form the Philox-4x32-10 counter of the packet's private
 * random stream: counter = {draw, 0, id_lo, id_hi}, key = seed. */
typedef struct {
  int64_t capacity;
  double *x, *y, *z;        /* position (cm)                       */
  double *Dx, *Dy, *Dz;     /* direction                           */
  double *t;                /* time (s)                            */
  double *e;                /* energy (erg)                        */
  double *nu;               /* frequency (Hz; MeV for gamma-rays)  */
  int32_t *ind;             /* zone                                */
  int32_t *type;            /* SGPU_PHOTON / SGPU_GAMMARAY         */
  int32_t *fate;            /* SGPU_MOVING ... SGPU_ABSORBED       */
  uint32_t *draw;           /* Philox draw counter                 */
  uint64_t *id;             /* Philox stream id                    */
} sgpu_packets;

/* Read-only opacity tables consumed by propagation. */
typedef struct {
  const double* abs;           /* [nz][n_nu] */
  const double* scat;          /* [nz][n_nu] (NULL when omit_scattering) */
  const double* emis_cdf;      /* [nz][n_nu] */
  const double* eps_imc;       /* [nz] */
  const double* compton_opac;  /* [nz] */
  const double* photoion_opac; /* [nz] */
} sgpu_tables;

/* Escaping-packet histogram (spectrum_array, src/transport/spectrum_array.cpp:36-83). */
typedef struct {
  sgpu_locate time_grid;
  sgpu_locate nu_grid;
  sgpu_locate mu_grid;
  sgpu_locate phi_grid;
  double* flux;             /* [n_t*n_nu*n_mu*n_phi] */
  double* click;            /* same                  */
} sgpu_spectrum;

/* Zone tallies (zone.h e_abs/e_rad/f*_rad/L_radio_*, transport::J_nu_). */
typedef struct {
  double* e_abs;
  double* J_nu;             /* [nz][n_nu] (or [nz][1] when store_Jnu = 0) */
  double* fx_rad;
  double* fy_rad;
  double* fz_rad;
  double* fr_rad;
  double* L_radio_dep;
  double* L_radio_emit;
  double* e_rad;
  sgpu_spectrum optical;
  sgpu_spectrum gamma;
} sgpu_tallies;

/* Transport parameters (transport members, transport.h:150-260). */
typedef struct {
  double  t_now;               /* transport::t_now_ at step start           */
  double  dt;                  /* step (1 in steady state)                  */
  uint64_t seed;               /* Philox key                                */
  int32_t steady_state;        /* transport_steady_iterate > 0              */
  int32_t radiative_eq;        /* transport_radiative_equilibrium          */
  int32_t boundary_in_reflect; /* transport_boundary_in_reflect             */
  int32_t boundary_out_reflect;/* transport_boundary_out_reflect            */
  int32_t store_Jnu;           /* transport_store_Jnu                       */
  int32_t omit_scattering;     /* opacity_no_scattering                     */
  int32_t precision;           /* SGPU_PREC_*                               */
  int32_t max_events;          /* 0 = unbounded; else packets still moving after
                                  this many events are left 'moving'          */
  int32_t use_ddmc;            /* transport_use_ddmc: 0 off, 2 discrete diffusion
                                  Monte Carlo (1D sphere, FP64 form only)     */
  double  ddmc_tau_threshold;  /* transport_ddmc_tau_threshold: zone optical
                                  depth above which a photon diffuses        */
} sgpu_transport_params;

/* Device counters of one propagation call (roofline inputs). */
typedef struct {
  uint64_t n_packets;
  uint64_t n_events;        /* loop iterations of propagate_monte_carlo */
  uint64_t n_binshift;      /* events whose distance was the nu-bin step */
  uint64_t n_boundary;      /* zone-boundary events (incl. bin steps)    */
  uint64_t n_scatter;       /* interaction events                        */
  uint64_t n_redist;        /* emissivity-CDF redistributions            */
  uint64_t n_escaped;
  uint64_t n_absorbed;
  uint64_t n_stopped;
  uint64_t n_moving;        /* left moving by max_events                 */
} sgpu_counters;

/* ------------------------------------------------------------------ */
/* context, memory, errors                                             */
/* ------------------------------------------------------------------ */
typedef struct sgpu_ctx sgpu_ctx;

/* Create a context on HIP device `device` with its own (default-priority)
 * stream.  Returns NULL on failure (then sgpu_last_error explains). */
sgpu_ctx* sgpu_create(int device);
void      sgpu_destroy(sgpu_ctx* ctx);
/* Make the context issue its work on an existing hipStream_t (passed as void*);
 * NULL selects the HIP null (default) stream.  The context starts on its own
 * stream, created blocking so that it is ordered with the null stream. */
int  sgpu_set_stream(sgpu_ctx* ctx, void* hip_stream);
void* sgpu_get_stream(sgpu_ctx* ctx);
int  sgpu_synchronize(sgpu_ctx* ctx);
/* Back to the context's own (created, blocking) stream after sgpu_set_stream. */
int  sgpu_use_own_stream(sgpu_ctx* ctx);

/* HIP graph of a whole step (no reference counterpart: the reference has no
 * GPU).  sgpu_graph_begin starts capturing the context's stream (a created
 * stream, not the null stream); the sgpu_* calls that follow are recorded,
 * not run -- they must make no host round trip (the *_dev entry points below
 * keep the packet counts on the device; sgpu_opacity_build with
 * zone_status NULL) and allocate nothing (every call runs once uncaptured
 * first, at the same sizes, so that its scratch exists).  sgpu_graph_end
 * ends the capture, updates the context's executable graph in place when the
 * topology matches the previous capture's (otherwise instantiates it anew)
 * and, with launch != 0, launches it.  Kernel arguments such as the step's
 * times are fixed when captured, so a time-dependent run captures every
 * step.  sgpu_graph_launch replays the last graph; sgpu_graph_nodes gives
 * its node count. */
int  sgpu_graph_begin(sgpu_ctx* ctx);
int  sgpu_graph_end(sgpu_ctx* ctx, int32_t launch);
int  sgpu_graph_launch(sgpu_ctx* ctx);
int64_t sgpu_graph_nodes(sgpu_ctx* ctx);
/* Copy the last error message of this thread into buf (NUL-terminated). */
int  sgpu_last_error(char* buf, size_t n);
int  sgpu_alloc(sgpu_ctx* ctx, void** ptr, size_t bytes);
int  sgpu_free(sgpu_ctx* ctx, void* ptr);
int  sgpu_memcpy_h2d(sgpu_ctx* ctx, void* dst, const void* src, size_t bytes);
int  sgpu_memcpy_d2h(sgpu_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Elapsed milliseconds of the most recent kernel of the named phase, timed
 * with HIP events on the context's stream ("opacity", "emit", "propagate",
 * "nlte").  Returns -1 when none was timed. */
double sgpu_last_kernel_ms(sgpu_ctx* ctx, const char* phase);
/* Enable (1) or disable (0) the per-phase HIP event timing above. */
int  sgpu_enable_timing(sgpu_ctx* ctx, int on);

/* ------------------------------------------------------------------ */
/* K1-K3: opacity / emissivity build                                   */
/* ------------------------------------------------------------------ */
/* For every zone: temperature clamp, radioactive decay of the composition,
 * LTE state solve by the reference's Brent iteration for n_e (only where
 * grey_opacity == 0), computeOpacity (grey, or electron scattering + fuzz
 * expansion), emissivity CDF, L_thermal, max-extinction clamp, Planck and
 * Rosseland means, gamma-ray coefficients and eps_imc.
 * Returns SGPU_OK, or the largest zone_status over zones (1 or 2) when some
 * zone's n_e solve failed (the tables are still written, as the reference
 * continues after a warning), or a negative error. */
int sgpu_opacity_build(sgpu_ctx* ctx,
                       const sgpu_locate* nu_grid,
                       const sgpu_zone_state* zones,
                       const sgpu_atoms* atoms,
                       const sgpu_opacity_params* params,
                       sgpu_opacity_out* out);

/* ------------------------------------------------------------------ */
/* K5: radioactive packet emission                                     */
/* ------------------------------------------------------------------ */
/* Emits n_emit packets (56Ni/56Co/48Cr chains) into packets[offset ..
 * offset+n_emit), stream ids id_base + q.  Adds the positron share to
 * tallies->L_radio_dep and writes tallies->L_radio_emit.  *n_out receives the
 * number emitted (0 when the grid's decay luminosity is zero). */
int sgpu_emit_radioactive(sgpu_ctx* ctx,
                          const sgpu_grid* grid,
                          const sgpu_locate* nu_grid,
                          const sgpu_zone_state* zones,
                          const sgpu_tables* tables,
                          const sgpu_transport_params* params,
                          int64_t n_emit, int32_t force_rproc, uint64_t id_base,
                          sgpu_packets* packets, int64_t offset,
                          sgpu_tallies* tallies, int64_t* n_out);

/* Inner boundary ("core") source (transport::setup_core_emission,
 * src/transport/transport_init.cpp:465-548; parameters core_radius,
 * core_temperature, core_luminosity, core_photon_frequency). */
typedef struct {
  double r_core;           /* cm; 0: packets start at the centre, isotropic     */
  double T_core;           /* K; 0 with L_core and r_core set: from
                              L = 4 pi r^2 sigma T^4; still 0: flat spectrum    */
  double L_core;           /* erg/s; 0: the blackbody's 4 pi^2 r^2 sum B_nu dnu */
  double core_frequency;   /* Hz; > 0: every packet at this frequency          */
} sgpu_core_source;

/* transport::emit_inner_source (src/transport/emission.cpp:370-480): n_emit
 * photon packets into packets[offset .. offset+n_emit), stream ids
 * id_base + q, on the sphere r_core (1 + 1e-10) at a uniform point with a
 * Lambertian (cos theta) outward direction -- or at the centre, isotropic,
 * for r_core = 0 -- energy L_core dt / n_emit, frequency from the core's
 * spectrum over the nu grid (B_nu(T_core) dnu at the bin centres, or
 * core_frequency), time t_now + u dt, Lorentz-transformed from the comoving
 * to the lab frame; zone from the grid.  As the reference, nothing is done
 * (*n_out = 0) when the packets do not fit the buffer.  The core spectrum
 * read from a file (core_spectrum_file) is not supported. */
int sgpu_emit_inner_source(sgpu_ctx* ctx,
                           const sgpu_grid* grid,
                           const sgpu_locate* nu_grid,
                           const sgpu_core_source* core,
                           const sgpu_transport_params* params,
                           int64_t n_emit, uint64_t id_base,
                           sgpu_packets* packets, int64_t offset,
                           int64_t* n_out);

/* Initial radiation field (transport::initialize_particles,
 * src/transport/emission.cpp:112-159, with create_isotropic_particle
 * :54-105): n_init photon packets into packets[offset .. offset+n_init),
 * stream ids id_base + q, the zone drawn from a CDF over e_rad[z] vol[z],
 * the position uniform in the zone's volume, the direction isotropic and the
 * frequency from B_nu(T_gas) dnu in the comoving frame, energy
 * sum(e_rad vol) / n_init, time params->t_now, then Lorentz-transformed to
 * the lab frame.  As the reference overwrites its emissivity, emis_cdf
 * [nz][n_nu] receives every zone's normalised blackbody CDF (the next
 * opacity build replaces it).  e_rad [nz] erg/cm^3.  As the reference,
 * nothing is done (*n_out = 0, emis_cdf kept) when n_init is 0 or the
 * packets do not fit the buffer. */
int sgpu_initialize_particles(sgpu_ctx* ctx,
                              const sgpu_grid* grid,
                              const sgpu_locate* nu_grid,
                              const sgpu_zone_state* zones,
                              const double* e_rad,
                              double* emis_cdf,
                              const sgpu_transport_params* params,
                              int64_t n_init, uint64_t id_base,
                              sgpu_packets* packets, int64_t offset,
                              int64_t* n_out);

/* ------------------------------------------------------------------ */
/* K4: propagation                                                     */
/* ------------------------------------------------------------------ */
/* Propagates packets [0, n) to the end of the step (or to escape / absorption
 * in steady state), tallies e_abs, J_nu, f_rad, L_radio_dep, and counts
 * escaping packets into the optical / gamma spectra at t_obs = t - x.D/c
 * (then sets t = t_obs, as transport::step does).  `counters` (device
 * pointer to an sgpu_counters, may be NULL) is accumulated. */
int sgpu_propagate(sgpu_ctx* ctx,
                   const sgpu_grid* grid,
                   const sgpu_locate* nu_grid,
                   const sgpu_tables* tables,
                   const sgpu_transport_params* params,
                   sgpu_packets* packets, int64_t n,
                   sgpu_tallies* tallies,
                   sgpu_counters* counters);

/* Removes escaped and absorbed packets, keeping order of the survivors
 * (src/transport/transport.cpp:164-194 keeps a different order; packets are
 * independent, so only the set matters).  `scratch` is an (n) packet-sized
 * staging buffer set (device), may alias nothing else.  *n_alive and
 * *n_escaped are host outputs. */
int sgpu_compact(sgpu_ctx* ctx, sgpu_packets* packets, int64_t n,
                 sgpu_packets* scratch, int64_t* n_alive, int64_t* n_escaped);

/* sgpu_compact without any device allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-block counts and totals live in a fixed device array of the library
 * (8 MiB + 256 B per GPU, held until the process ends), so n is limited to
 * 2^28 = 268,435,456 packets (SGPU_ERR_SIZE beyond).  packets and scr are
 * caller-owned as for sgpu_compact, and the results are identical to it.
 * Calls that overlap on one GPU share that array and race. */
int sgpu_compact_noalloc(sgpu_ctx* ctx, sgpu_packets* packets, int64_t n, sgpu_packets* scr,
                         int64_t* n_alive, int64_t* n_escaped);

/* Device-resident packet counts (no reference counterpart), for a step
 * enqueued without host round trips, e.g. captured in a HIP graph:
 *   sgpu_emit_radioactive_dev: as sgpu_emit_radioactive with the offset read
 *     from the device int64 *d_count, which it then advances by n_emit when
 *     packets were emitted (non-zero decay luminosity, room in the buffer:
 *     otherwise nothing is emitted, as the reference's "Out of particle
 *     space" case);
 *   sgpu_propagate_dev: as sgpu_propagate over packets [0, *d_n), the
 *     launches sized for the buffer's capacity;
 *   sgpu_compact_dev: as sgpu_compact over [0, *d_count), leaving the
 *     survivor count in *d_count and the escapes in *d_escaped (device; may
 *     be NULL); scr must be at least as large as packets. */
int sgpu_emit_radioactive_dev(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                              const sgpu_zone_state* zones, const sgpu_tables* tables,
                              const sgpu_transport_params* params, int64_t n_emit, int32_t force_rproc,
                              uint64_t id_base, sgpu_packets* packets, int64_t* d_count, sgpu_tallies* tallies);
int sgpu_propagate_dev(sgpu_ctx* ctx, const sgpu_grid* grid, const sgpu_locate* nu_grid,
                       const sgpu_tables* tables, const sgpu_transport_params* params,
                       sgpu_packets* packets, const int64_t* d_n, sgpu_tallies* tallies,
                       sgpu_counters* counters);
int sgpu_compact_dev(sgpu_ctx* ctx, sgpu_packets* packets, int64_t* d_count, sgpu_packets* scr,
                     int64_t* d_escaped);

/* Throughput calibration of the per-event J_nu tally (no reference
 * counterpart): `adds` FP64 atomic additions of 1.0 at hash-scattered
 * indices of buf[0, n) (device), launched on the context's stream and timed
 * with HIP events; *ms receives the kernel time.  buf then holds the counts
 * (they sum to `adds`).  The propagation does one such addition per event
 * into J_nu[zone][bin]. */
int sgpu_atomic_rate(sgpu_ctx* ctx, double* buf, int64_t n, int64_t adds, double* ms);

/* Zero the zone and spectrum tallies (transport::wipe_radiation; the spectra
 * are wiped only when wipe_spectra != 0). */
int sgpu_wipe_radiation(sgpu_ctx* ctx, int32_t n_zones, int32_t n_nu,
                        sgpu_tallies* tallies, int32_t wipe_spectra);

/* reduce_radiation normalisation (transport_mpi.cpp:415-441). */
int sgpu_normalize_tallies(sgpu_ctx* ctx, const sgpu_grid* grid,
                           const sgpu_locate* nu_grid,
                           const sgpu_transport_params* params,
                           sgpu_tallies* tallies);

/* ------------------------------------------------------------------ */
/* K8: NLTE level populations                                          */
/* ------------------------------------------------------------------ */
/* One system per call, FP64 partial-pivot LU in device memory: the
 * drop-in for AtomicSpecies::solve_nlte.  rates[n*n] (row = from-level,
 * already multiplied by the LTE population of the source level, as
 * set_rates leaves them) and lte[n] are device pointers; the solver builds
 * M_ii = -sum_j R_ij, M_ij = R_ji, last row = lte, rhs e_{n-1}, factors and
 * solves; b[n] receives the departure coefficients and pop[n] (may be NULL)
 * max(b*lte, 1e-30).  work must hold n*n + n doubles (device). */
int sgpu_nlte_solve(sgpu_ctx* ctx, int32_t n, const double* rates,
                    const double* lte, double* b, double* pop, double* work);

/* All systems in one persistent launch.  System s has size n[s]; its rates
 * start at rates + rate_off[s] (n[s]^2 doubles, row-major) and its lte/b/pop
 * at vec_off[s].  precision: SGPU_NLTE_FP64 or SGPU_NLTE_FP32_REFINE.
 * FP32 mode: refinement stops when max|dx|/max|x| <= tol, or, once the
 * corrections no longer halve from one iteration to the next (the FP64
 * residual's rounding floor), when max|dx|/max|x| <= 16 tol; at most 12
 * corrections.
 * iters[s] (device, may be NULL) receives the refinement iterations used, or
 * -1 when the FP32 solve did not converge and the system was re-solved in FP64.
 * n[], rate_off[], vec_off[] are device arrays. */
int sgpu_nlte_batched(sgpu_ctx* ctx, int32_t nsys, const int32_t* n,
                      const int64_t* rate_off, const int64_t* vec_off,
                      const double* rates, const double* lte,
                      double* b, double* pop,
                      int32_t precision, double tol, int32_t* iters);

/* Rate assembly of AtomicSpecies::set_rates (AtomicSpecies_state.cpp:294-432)
 * for a batch of (zone, atom) systems, in the layout sgpu_nlte_batched reads.
 * System s is atom sys_atom[s] at electron density ne[s] and temperature
 * T[s]; its levels are the atom's levels (local index i = global level
 * - lev_off[atom]).  Per-system inputs are indexed by GLOBAL line / level:
 *   line_J [s*n_blines + l]   mean intensity J-bar of line l
 *   Pic, Rci [s*n_levels + k] photoionisation and recombination rate
 *                             coefficients of level k (calculate_radiative_rates)
 *   lev_lte [s*n_levels + k]  LTE population fraction (solve_lte)
 * Terms, in the reference's order: B_ul J + A and B_lu J per line; van
 * Regemorter collisions (f_lu floored at 1e-3, zero upward rate beyond
 * zeta = 700); collisional ionisation / recombination; R_ci n_e and P_ic
 * (R_ci of E = 0 levels dropped when no_ground_recomb); every rate out of
 * level i times lev_lte[i].  Outputs rates[rate_off[s] + i*n + j] and
 * lte[vec_off[s] + i]; n[s] must equal the atom's level count.  Each matrix
 * entry is summed by one thread in the reference's term order. */
int sgpu_nlte_set_rates(sgpu_ctx* ctx, const sgpu_atoms* atoms, int32_t nsys,
                        const int32_t* sys_atom, const double* ne, const double* T,
                        const double* line_J, const double* Pic, const double* Rci,
                        const double* lev_lte, int32_t use_collisions, int32_t no_ground_recomb,
                        const int32_t* n, const int64_t* rate_off, const int64_t* vec_off,
                        double* rates, double* lte);

/* transport::set_opacity with the NLTE gas state (transport_opacity.cpp:
 * 133-136 -> GasState::solve_state(J_nu) under opacity_use_nlte,
 * GasState.cpp:208-263): the opacity build of sgpu_opacity_build, except that
 * at every evaluation of the n_e Brent iteration the atoms flagged in
 * atom_nlte (HOST array [n_atoms], opacity_atoms_in_nlte) are solved by
 * AtomicSpecies::solve_nlte (AtomicSpecies_state.cpp:131-196: LTE at n_e,
 * set_rates with the radiative rates computed once per zone from J_nu
 * [nz][n_nu] (device), LU, populations) and the others by solve_lte.  The
 * zone keeps the state of its last evaluation, as the reference's atoms do.
 * lu_precision: SGPU_NLTE_FP64 or SGPU_NLTE_FP32_REFINE for the batched
 * solves; use_collisions: opacity_use_collisions_nlte; the line Doppler
 * width is params->line_velocity_width as for sgpu_nlte_radiative_rates.
 * On the first step (params->first_step)
 * or with no atom flagged it is sgpu_opacity_build.  *n_evals (may be NULL)
 * receives the number of n_e evaluation passes. */
int sgpu_opacity_build_nlte(sgpu_ctx* ctx, const sgpu_locate* nu_grid, const sgpu_zone_state* zones,
                            const sgpu_atoms* atoms, const sgpu_opacity_params* params,
                            const double* J_nu, const int32_t* atom_nlte, int32_t use_collisions,
                            int32_t lu_precision, sgpu_opacity_out* out, int32_t* n_evals);

/* AtomicSpecies::calculate_radiative_rates (AtomicSpecies_state.cpp:202-287)
 * for a batch of (zone, atom) systems: from the zone's mean intensity
 * J_nu[sys_zone[s]][n_nu] (the transport tally) and gas temperature T[s],
 * the photoionisation and recombination rate coefficients of every level
 * with a continuum (hydrogenic cross-sections, AtomicData.cpp:713-746),
 * summed over bins 1..n_nu-1, and each line's J-bar = the trapezoid over
 * x in [-5, 5] (step 0.05) of J(nu_0 + x nu_D) phi_Voigt(x, A / (4 pi nu_D)),
 * nu_D = nu_0 sqrt(2 k T / (m_p elem_A[atom])) / c (or line_velocity_width
 * when > 0; GasState.cpp:218-220).  Outputs are indexed as
 * sgpu_nlte_set_rates reads them: Pic / Rci [s*n_levels + level],
 * line_J [s*n_blines + line]; only the system's atom's entries are written. */
int sgpu_nlte_radiative_rates(sgpu_ctx* ctx, const sgpu_atoms* atoms, const sgpu_locate* nu_grid,
                              int32_t nsys, const int32_t* sys_atom, const int32_t* sys_zone,
                              const double* T, const int32_t* elem_A, double line_velocity_width,
                              const double* J_nu, double* Pic, double* Rci, double* line_J);

/* End of AtomicSpecies::solve_nlte (AtomicSpecies_state.cpp:178-193):
 * lev_n = max(b * lev_lte, 1e-30) per level and ion_frac[s*n_ions + ion] =
 * sum of lev_n over the ion's levels (global ion index; other atoms' entries
 * are left untouched). */
int sgpu_nlte_populations(sgpu_ctx* ctx, const sgpu_atoms* atoms, int32_t nsys,
                          const int32_t* sys_atom, const int64_t* vec_off,
                          const double* b, const double* lte, double* lev_n, double* ion_frac);

/* Choose launch shapes for sgpu_nlte_batched by timing candidates on the
 * given batch (project-written autotuner): workgroup size {128, 256, 512} x
 * LU panel width {8, 16} x workgroups per CU {1, 2, 4}, for each
 * precision, keeping only shapes whose workgroups can share a CU (LDS and
 * registers).  The choice is kept in the context and used by later
 * sgpu_nlte_batched calls. */
int sgpu_nlte_autotune(sgpu_ctx* ctx, int32_t nsys, const int32_t* n,
                       const int64_t* rate_off, const int64_t* vec_off,
                       const double* rates, const double* lte,
                       double* b, double* pop, int32_t* iters,
                       char* report, size_t report_len);

/* Pin the launch shape of sgpu_nlte_batched for one precision
 * (SGPU_NLTE_FP64 or SGPU_NLTE_FP32_REFINE) instead of autotuning:
 * threads a multiple of 64 in [128, 512], panel 8 or 16, wg_per_cu in [1, 16]
 * (project-written; no reference counterpart). */
int sgpu_nlte_set_tuning(sgpu_ctx* ctx, int32_t precision, int32_t threads, int32_t panel,
                         int32_t wg_per_cu);

/* ------------------------------------------------------------------ */
/* radiative-equilibrium gas temperature                               */
/* ------------------------------------------------------------------ */
/* Walker alias tables of the zones' emissivity distributions -- the
 * frequency sampler of the FP32 propagation form, standing in for the
 * binary search of transport::sample_photon_frequency over emissivity_[i]
 * (emission.cpp:27-46).  Column k of zone z keeps bin k with probability
 * q and otherwise samples bin alias; the distribution is p_k = cdf_k -
 * cdf_{k-1} over the normalised CDF emis_cdf [n_zones][n_nu] (device), with
 * the last bin taking the remainder below 1.  table [n_zones][n_nu] (device).
 * n_nu <= 4096. */
typedef struct {
  float q;
  int32_t alias;
} sgpu_alias_entry;
int sgpu_alias_tables(sgpu_ctx* ctx, const double* emis_cdf, int32_t n_zones, int32_t n_nu,
                      sgpu_alias_entry* table);

/* transport::solve_eq_temperature (solve_equilibrium.cpp:66-94): per zone,
 * Brent on [temp_min, temp_max] (tolerance 1e-2 K, EPS 3e-8, at most 100
 * iterations, then 0.0 as the reference returns) for the T at which the
 * Kirchhoff emission sum_nu 4 pi abs_opacity[z][nu] B_nu(T) dnu (one bin:
 * 4 pi abs sigma T^4 / pi) equals the e_abs tally -- the LTE balance with
 * the step's opacities (solve_flag 0, :112-168).  With set_Tgas_to_Trad,
 * T = (e_rad / a)^(1/4).  T_gas[nz] is written; n_eval[nz] (may be NULL)
 * receives the balance evaluations of each zone. */
int sgpu_solve_eq_temperature(sgpu_ctx* ctx, const sgpu_locate* nu_grid, int32_t n_zones,
                              const double* abs_opacity, const double* e_abs, const double* e_rad,
                              double temp_min, double temp_max, int32_t set_Tgas_to_Trad,
                              double* T_gas, int32_t* n_eval);

/* ------------------------------------------------------------------ */
/* homologous expansion                                                */
/* ------------------------------------------------------------------ */
/* rho /= e^3, grid edges and volumes scaled by e (host-side descriptor
 * grid->t_now is NOT changed: the caller owns it).  rho, edges and vol are
 * device arrays; e = (t+dt)/t. */
int sgpu_homologous_step(sgpu_ctx* ctx, sgpu_grid* grid, double* rho,
                         double* r_out, double* x_out, double* z_out,
                         double* vol, double e);

/* ------------------------------------------------------------------ */
/* differentiable transport (continuous-absorption estimator)          */
/* ------------------------------------------------------------------ */
/* Forward pass: 1D grid, steady state.  kappa_abs[nz][n_nu] (FP32, cm^2/g)
 * times rho gives the absorption coefficient; scattering scat[nz][n_nu]
 * (cm^-1, FP64) is sampled discretely; absorption acts as a weight
 * w = exp(-sum kappa rho dshift ds).  Escaping weight*energy is binned into
 * F[n_nu_spec] (FP64).  rec_bin[n] / rec_w[n] (device, may be NULL) keep each
 * packet's spectrum bin (-1 when not counted) and final weight for the
 * adjoint replay. */
int sgpu_adjoint_forward(sgpu_ctx* ctx, const sgpu_grid* grid,
                         const sgpu_locate* nu_grid, const sgpu_locate* spec_grid,
                         const float* kappa_abs, const double* rho,
                         const double* scat, const sgpu_transport_params* params,
                         const sgpu_packets* packets, int64_t n,
                         double* F, int32_t* rec_bin, double* rec_w);

/* Adjoint replay: given dL/dF[n_nu_spec] and the forward records, replays
 * every packet path from its Philox counters and accumulates
 * dL/dkappa_abs[nz][n_nu] (FP64 accumulation, device). */
int sgpu_adjoint_backward(sgpu_ctx* ctx, const sgpu_grid* grid,
                          const sgpu_locate* nu_grid, const sgpu_locate* spec_grid,
                          const float* kappa_abs, const double* rho,
                          const double* scat, const sgpu_transport_params* params,
                          const sgpu_packets* packets, int64_t n,
                          const double* dL_dF, const int32_t* rec_bin,
                          const double* rec_w, double* dL_dkappa);

#ifdef __cplusplus
}
#endif

#endif /* SEDONA_GPU_H */
