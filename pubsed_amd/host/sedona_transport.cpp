// sedona_transport.cpp -- host C++ drop-in for Sedona's transport step (see
// sedona_transport.hpp).  Only C-ABI calls; no HIP or torch types.
#include "sedona_transport.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace sedona_gpu {

namespace {
constexpr double PI_REF = 3.14159;   // src/utils/physical_constants.h:9
constexpr double A_RAD_REF = 7.5657e-15;   // radiation constant, physical_constants.h:17

// locate_array::log_init with a spacing nudged until it yields exactly n bins
std::vector<double> log_edges(double lo, double hi, int n, double* del_out)
{
  double d0 = std::pow(hi / lo, 1.0 / n) - 1.0;
  for (int k = 1; k <= 40; ++k) {
    double d = d0 * (1.0 + 1e-12 * k);
    std::vector<double> x;
    double v = lo + lo * d;
    while (v < hi && (int)x.size() <= n) { x.push_back(v); v = v + v * d; }
    x.push_back(hi);
    if ((int)x.size() == n) { *del_out = d; return x; }
  }
  *del_out = 0;
  return {};
}

// locate_array::init(start, stop, double del) (locate_array.cpp:25-46) and
// init(start, stop, int n) (:88-108); both give one catch-all bin (type none)
// for an empty range
sgpu_locate lin_grid(double start, double stop, double del, int n, HostGrid* h)
{
  const bool none = n < 0 ? start >= stop : (start == stop || n == 0);
  if (none) {
    h->min = -INFINITY;
    h->x.assign(1, INFINITY);
    return {1, SGPU_LOC_NONE, -INFINITY, 0.0, nullptr};
  }
  if (n < 0) n = std::max((int)std::ceil((stop - start) / del), 1);
  else del = (stop - start) / (double)n;
  h->min = start;
  h->x.resize(n);
  for (int i = 0; i < n - 1; i++) h->x[i] = start + (i + 1) * del;
  h->x[n - 1] = stop;
  return {n, SGPU_LOC_LIN, start, del, nullptr};
}
}  // namespace

transport::~transport()
{
  for (void* p : owned_) sgpu_free(ctx_, p);
  if (ctx_) sgpu_destroy(ctx_);
}

int transport::fail(int rc, const char* what)
{
  char buf[1024];
  sgpu_last_error(buf, sizeof(buf));
  err_ = std::string(what) + ": " + buf;
  return rc;
}

void* transport::dalloc(size_t bytes)
{
  void* p = nullptr;
  if (sgpu_alloc(ctx_, &p, bytes) != SGPU_OK) return nullptr;
  owned_.push_back(p);
  return p;
}

template <class T>
T* transport::upload(const std::vector<T>& v)
{
  T* d = (T*)dalloc(sizeof(T) * (v.empty() ? 1 : v.size()));
  if (d && !v.empty()) sgpu_memcpy_h2d(ctx_, d, v.data(), sizeof(T) * v.size());
  return d;
}

std::vector<double> transport::download(const double* d, size_t n) const
{
  std::vector<double> h(n);
  if (n) sgpu_memcpy_d2h(ctx_, h.data(), d, sizeof(double) * n);
  return h;
}

int transport::init(const Params& p, const Grid1D& g, const AtomicData* atoms, int device)
{
  p_ = p;
  g_ = g;
  ctx_ = sgpu_create(device);
  if (!ctx_) return fail(SGPU_ERR_HIP, "sgpu_create");
  nz_ = g.n_zones();
  ne_ = (int)g.elems_Z.size();
  t_now_ = g.t_now;
  double del = 0;
  nu_edges_ = log_edges(p.transport_nu_grid[0], p.transport_nu_grid[1], p.n_nu, &del);
  if ((int)nu_edges_.size() != p.n_nu) return fail(SGPU_ERR_ARG, "frequency grid");
  nn_ = p.n_nu;
  nu_ = {nn_, SGPU_LOC_LOG, p.transport_nu_grid[0], del, upload(nu_edges_)};

  // grid: shell volumes 4/3 pi (r^3 - r0^3) with the reference's pi
  std::vector<double> vol(nz_);
  for (int i = 0; i < nz_; ++i) {
    double r0 = i == 0 ? g.r_min : g.r_out[i - 1];
    vol[i] = 4.0 * PI_REF / 3.0 * (g.r_out[i] * g.r_out[i] * g.r_out[i] - r0 * r0 * r0);
  }
  d_r_out_ = upload(g.r_out);
  d_vol_ = upload(vol);
  grid_ = {};
  grid_.kind = SGPU_GRID_1D_SPHERE;
  grid_.n_zones = nz_;
  grid_.t_now = g.t_now;
  grid_.r_core = p.core_radius;
  grid_.r_out = {nz_, SGPU_LOC_FLEX, g.r_min, 0.0, d_r_out_};
  grid_.vol = d_vol_;

  d_rho_ = upload(g.rho);
  std::vector<double> grey(nz_, p.opacity_grey_opacity);
  zs_ = {nz_, ne_, d_rho_, upload(g.T_gas), upload(g.X), nullptr, upload(grey),
         upload(g.elems_Z), upload(g.elems_A)};
  have_atoms_ = atoms != nullptr;
  if (have_atoms_) {
    at_.n_atoms = (int)atoms->lev_off.size() - 1;
    at_.max_levels = atoms->max_levels;
    at_.n_levels = atoms->lev_off.back();
    at_.n_ions = atoms->ion_off.back();
    at_.lev_off = upload(atoms->lev_off);
    at_.ion_off = upload(atoms->ion_off);
    at_.lev_E = upload(atoms->lev_E);
    at_.lev_g = upload(atoms->lev_g);
    at_.lev_ion = upload(atoms->lev_ion);
    at_.ion_chi = upload(atoms->ion_chi);
    at_.atom_zero_eps = upload(atoms->atom_zero_eps);
    at_.n_lines = (int64_t)atoms->line_nu.size();
    at_.line_nu = upload(atoms->line_nu);
    at_.line_gf = upload(atoms->line_gf);
    at_.line_El = upload(atoms->line_El);
    at_.line_ion = upload(atoms->line_ion);
    at_.line_bin = upload(atoms->line_bin);
    at_.line_bin_off = upload(atoms->line_bin_off);
    at_.n_blines = (int64_t)atoms->bline_nu.size();
    if (at_.n_blines > 0) {
      at_.bline_l = upload(atoms->bline_l);
      at_.bline_u = upload(atoms->bline_u);
      at_.bline_nu = upload(atoms->bline_nu);
      at_.bline_f = upload(atoms->bline_f);
      at_.bline_A = upload(atoms->bline_A);
      at_.bline_Bul = upload(atoms->bline_Bul);
      at_.bline_Blu = upload(atoms->bline_Blu);
      at_.bline_bin = upload(atoms->bline_bin);
      at_.bline_bin_off = upload(atoms->bline_bin_off);
    }
  }
  const size_t T = (size_t)nz_ * nn_;
  auto z = [&](size_t n) { return (double*)dalloc(sizeof(double) * n); };
  out_ = {};
  out_.abs = z(T); out_.scat = z(T); out_.emis_cdf = z(T);
  out_.L_thermal = z(nz_); out_.planck_mean = z(nz_); out_.rosseland_mean = z(nz_);
  out_.n_elec = z(nz_); out_.compton_opac = z(nz_); out_.photoion_opac = z(nz_); out_.eps_imc = z(nz_);
  out_.zone_status = (int32_t*)dalloc(sizeof(int32_t) * nz_);
  tab_ = {out_.abs, out_.scat, out_.emis_cdf, out_.eps_imc, out_.compton_opac, out_.photoion_opac};

  // spectrum_array::init (spectrum_array.cpp:36-83) as transport_init.cpp:
  // 166-173 calls it: spectrum_time_grid, the transport nu grid, n_mu bins on
  // [-1, 1], n_phi on [0, 2 pi]; the gamma spectrum on the same grids
  sgpu_locate tg = lin_grid(p.spectrum_time_grid[0], p.spectrum_time_grid[1], p.spectrum_time_grid[2], -1, &spec_t_);
  sgpu_locate mg = lin_grid(-1.0, 1.0, 0.0, p.spectrum_n_mu, &spec_mu_);
  sgpu_locate pg = lin_grid(0.0, 2 * PI_REF, 0.0, p.spectrum_n_phi, &spec_phi_);
  tg.x = upload(spec_t_.x);
  mg.x = upload(spec_mu_.x);
  pg.x = upload(spec_phi_.x);
  n_spec_ = (size_t)tg.n * nn_ * mg.n * pg.n;
  ta_ = {};
  ta_.e_abs = z(nz_); ta_.J_nu = z(T); ta_.fx_rad = z(nz_); ta_.fy_rad = z(nz_); ta_.fz_rad = z(nz_);
  ta_.fr_rad = z(nz_); ta_.L_radio_dep = z(nz_); ta_.L_radio_emit = z(nz_); ta_.e_rad = z(nz_);
  ta_.optical = {tg, nu_, mg, pg, z(n_spec_), z(n_spec_)};
  ta_.gamma = {tg, nu_, mg, pg, z(n_spec_), z(n_spec_)};

  const int64_t cap = p.particles_max_total;
  for (sgpu_packets* q : {&pk_, &scr_}) {
    q->capacity = cap;
    q->x = z(cap); q->y = z(cap); q->z = z(cap);
    q->Dx = z(cap); q->Dy = z(cap); q->Dz = z(cap);
    q->t = z(cap); q->e = z(cap); q->nu = z(cap);
    q->ind = (int32_t*)dalloc(sizeof(int32_t) * cap);
    q->type = (int32_t*)dalloc(sizeof(int32_t) * cap);
    q->fate = (int32_t*)dalloc(sizeof(int32_t) * cap);
    q->draw = (uint32_t*)dalloc(sizeof(uint32_t) * cap);
    q->id = (uint64_t*)dalloc(sizeof(uint64_t) * cap);
    if (!q->id) return fail(SGPU_ERR_HIP, "packet buffers");
  }
  if (sgpu_wipe_radiation(ctx_, nz_, nn_, &ta_, 1) != SGPU_OK) return fail(SGPU_ERR_HIP, "wipe");
  // transport_init.cpp:376-379: the initial radiation field from the zones'
  // e_rad (a T_gas^4 unless the model gives it, grid_1D_sphere.cpp:130-134, 354)
  if (p.particles_n_initialize > 0) {
    std::vector<double> er = g.e_rad;
    if (er.empty())
      for (int i = 0; i < nz_; ++i) er.push_back(A_RAD_REF * std::pow(g.T_gas[i], 4.));
    if ((int)er.size() != nz_) return fail(SGPU_ERR_ARG, "e_rad needs one value per zone");
    if (sgpu_memcpy_h2d(ctx_, ta_.e_rad, er.data(), sizeof(double) * nz_) != SGPU_OK)
      return fail(SGPU_ERR_HIP, "e_rad upload");
    sgpu_transport_params tp{};
    tp.t_now = t_now_;
    tp.seed = p.transport_rng_seed;
    int64_t got = 0;
    // stream ids above every emission's (step_index << 32 | q)
    int rc = sgpu_initialize_particles(ctx_, &grid_, &nu_, &zs_, ta_.e_rad, out_.emis_cdf, &tp,
                                       p.particles_n_initialize, (uint64_t)0xFFFFFFFFull << 32, &pk_, 0, &got);
    if (rc < 0) return fail(rc, "initialize_particles");
    n_particles_ = got;
  }
  return SGPU_OK;
}

int transport::set_opacity(double dt)
{
  op_ = {};
  op_.t_now = t_now_;
  op_.dt = dt;
  op_.epsilon = p_.opacity_epsilon;
  op_.temp_min = p_.limits_temp_min;
  op_.temp_max = p_.limits_temp_max;
  op_.maximum_opacity = p_.opacity_maximum_opacity;
  op_.fleck_alpha = p_.transport_fleck_alpha;
  op_.use_electron_scattering = p_.opacity_electron_scattering;
  op_.use_fuzz_expansion = p_.opacity_fuzz_expansion;
  op_.use_line_expansion = p_.opacity_line_expansion;
  op_.use_free_free = p_.opacity_free_free;
  op_.use_bound_free = p_.opacity_bound_free;
  op_.no_ground_recomb = p_.opacity_no_ground_recomb;
  op_.use_bound_bound = p_.opacity_bound_bound;
  op_.radiative_eq = p_.transport_radiative_equilibrium;
  op_.first_step = first_step_;
  op_.line_velocity_width = p_.line_velocity_width;
  op_.minimum_extinction = p_.opacity_minimum_extinction;
  int rc;
  // transport_opacity.cpp:30-36, 133-136: the NLTE gas state from the J_nu of
  // the previous step (LTE on the first step, inside the call)
  if (p_.opacity_use_nlte && have_atoms_ && p_.transport_store_Jnu &&
      (int)p_.opacity_atoms_in_nlte.size() == at_.n_atoms)
    rc = sgpu_opacity_build_nlte(ctx_, &nu_, &zs_, &at_, &op_, ta_.J_nu, p_.opacity_atoms_in_nlte.data(),
                                 p_.opacity_use_collisions_nlte, SGPU_NLTE_FP64, &out_, nullptr);
  else
    rc = sgpu_opacity_build(ctx_, &nu_, &zs_, have_atoms_ ? &at_ : nullptr, &op_, &out_);
  if (rc < 0) return fail(rc, "set_opacity");
  return SGPU_OK;   // positive codes are the reference's solve warnings
}

int transport::wipe_radiation()
{
  int rc = sgpu_wipe_radiation(ctx_, nz_, nn_, &ta_, p_.transport_steady_iterate > 0);
  return rc < 0 ? fail(rc, "wipe_radiation") : SGPU_OK;
}

int transport::emit_particles(double dt)
{
  int64_t n = p_.particles_n_emit_radioactive;
  if (n_particles_ + n > p_.particles_max_total) n = 0;   // "Out of particle space"
  int64_t got = 0;
  int rc = sgpu_emit_radioactive(ctx_, &grid_, &nu_, &zs_, &tab_, &tp_, n, 0, step_index_ << 32, &pk_,
                                 n_particles_, &ta_, &got);
  if (rc < 0) return fail(rc, "emit_particles");
  n_particles_ += got;
  // emission.cpp:15-22: the inner boundary source after the radioactive one
  if (p_.core_n_emit > 0) {
    const sgpu_core_source core = {p_.core_radius, p_.core_temperature, p_.core_luminosity,
                                   p_.core_photon_frequency};
    rc = sgpu_emit_inner_source(ctx_, &grid_, &nu_, &core, &tp_, p_.core_n_emit, (step_index_ << 32) | (1ull << 31),
                                &pk_, n_particles_, &got);
    if (rc < 0) return fail(rc, "emit_inner_source");
    n_particles_ += got;
  }
  return SGPU_OK;
}

int transport::propagate_all(double dt)
{
  int rc = sgpu_propagate(ctx_, &grid_, &nu_, &tab_, &tp_, &pk_, n_particles_, &ta_, nullptr);
  return rc < 0 ? fail(rc, "propagate") : SGPU_OK;
}

int transport::clean_up_particle_vector()
{
  int64_t alive = 0, esc = 0;
  int rc = sgpu_compact(ctx_, &pk_, n_particles_, &scr_, &alive, &esc);
  if (rc < 0) return fail(rc, "clean_up_particle_vector");
  n_particles_ = alive;
  n_escaped_ = esc;
  return SGPU_OK;
}

int transport::reduce_radiation(double dt)
{
  int rc = sgpu_normalize_tallies(ctx_, &grid_, &nu_, &tp_, &ta_);
  return rc < 0 ? fail(rc, "reduce_radiation") : SGPU_OK;
}

int transport::step(double dt)
{
  if (p_.transport_steady_iterate > 0) dt = 1;
  tp_ = {};
  tp_.t_now = t_now_;
  tp_.dt = dt;
  tp_.seed = p_.transport_rng_seed;
  tp_.steady_state = p_.transport_steady_iterate > 0;
  tp_.radiative_eq = p_.transport_radiative_equilibrium;
  tp_.boundary_in_reflect = p_.transport_boundary_in_reflect;
  tp_.boundary_out_reflect = p_.transport_boundary_out_reflect;
  tp_.store_Jnu = p_.transport_store_Jnu;
  tp_.precision = p_.precision;
  tp_.use_ddmc = p_.transport_use_ddmc;
  tp_.ddmc_tau_threshold = p_.transport_ddmc_tau_threshold;
  int rc;
  if ((rc = set_opacity(dt))) return rc;
  if ((rc = wipe_radiation())) return rc;
  if ((rc = emit_particles(dt))) return rc;
  if ((rc = propagate_all(dt))) return rc;
  if ((rc = clean_up_particle_vector())) return rc;
  if ((rc = reduce_radiation(dt))) return rc;
  // transport.cpp:126-131: the radiative-equilibrium temperature for the
  // next step (solve_equilibrium.cpp:66-94); the zone temperatures are
  // device data of the gas state
  if (p_.transport_radiative_equilibrium == 1 && p_.transport_fix_Tgas_during_transport == 0) {
    rc = sgpu_solve_eq_temperature(ctx_, &nu_, nz_, out_.abs, ta_.e_abs, ta_.e_rad, p_.limits_temp_min,
                                   p_.limits_temp_max, p_.transport_set_Tgas_to_Trad, (double*)zs_.T_gas,
                                   nullptr);
    if (rc < 0) return fail(rc, "solve_eq_temperature");
  }
  if (sgpu_synchronize(ctx_) != SGPU_OK) return fail(SGPU_ERR_HIP, "synchronize");
  if (!tp_.steady_state) t_now_ += dt;
  first_step_ = 0;
  ++step_index_;
  return SGPU_OK;
}

std::vector<double> transport::J_nu() const { return download(ta_.J_nu, (size_t)nz_ * nn_); }
std::vector<double> transport::e_rad() const { return download(ta_.e_rad, nz_); }
std::vector<double> transport::e_abs() const { return download(ta_.e_abs, nz_); }
std::vector<double> transport::L_radio_dep() const { return download(ta_.L_radio_dep, nz_); }
std::vector<double> transport::optical_spectrum() const { return download(ta_.optical.flux, n_spec_); }
std::vector<double> transport::optical_clicks() const { return download(ta_.optical.click, n_spec_); }
std::vector<double> transport::abs_opacity() const { return download(out_.abs, (size_t)nz_ * nn_); }
std::vector<double> transport::T_gas() const { return download(zs_.T_gas, nz_); }
std::vector<double> transport::n_elec() const { return download(out_.n_elec, nz_); }

int transport::output_spectrum(int it)
{
  // the spectrum grids set up in init
  HostGrid ng{p_.transport_nu_grid[0], nu_edges_};
  const std::string base = it < 0 ? std::string("_final") : "_" + std::to_string(it);
  if (p_.spectrum_suppress_txt) return SGPU_OK;
  if (!p_.spectrum_name.empty() &&
      write_spectrum_dat(p_.spectrum_name + base, spec_t_, ng, spec_mu_, spec_phi_,
                         download(ta_.optical.flux, n_spec_), download(ta_.optical.click, n_spec_)))
    return fail(SGPU_ERR_ARG, "output_spectrum: cannot write the optical spectrum");
  if (!p_.gamma_name.empty() &&
      write_spectrum_dat(p_.gamma_name + base, spec_t_, ng, spec_mu_, spec_phi_,
                         download(ta_.gamma.flux, n_spec_), download(ta_.gamma.click, n_spec_)))
    return fail(SGPU_ERR_ARG, "output_spectrum: cannot write the gamma spectrum");
  return SGPU_OK;
}

namespace {
// value of a numeric key in the flat meta.json (the first number after the
// key's colon; an array's first element); dflt for a missing key or null
double meta_number(const std::string& js, const char* key, double dflt)
{
  const std::string k = std::string("\"") + key + "\"";
  size_t p = js.find(k);
  if (p == std::string::npos) return dflt;
  p = js.find(':', p + k.size());
  if (p == std::string::npos) return dflt;
  const char* c = js.c_str() + p + 1;
  while (*c == ' ' || *c == '\n' || *c == '[') ++c;
  char* end;
  const double v = strtod(c, &end);
  return end == c ? dflt : v;
}

bool make_dir(const std::string& d) { return mkdir(d.c_str(), 0755) == 0 || errno == EEXIST; }
}  // namespace

int transport::write_checkpoint(const std::string& dir)
{
  if (!make_dir(dir) || !make_dir(dir + "/particles") || !make_dir(dir + "/state"))
    return fail(SGPU_ERR_ARG, "write_checkpoint: cannot create the directory");
  const size_t n = (size_t)n_particles_;
  auto d2h = [&](void* h, const void* d, size_t bytes) { return bytes ? sgpu_memcpy_d2h(ctx_, h, d, bytes) : 0; };
  std::vector<double> c[9];
  const double* src[9] = {pk_.x, pk_.y, pk_.z, pk_.Dx, pk_.Dy, pk_.Dz, pk_.t, pk_.e, pk_.nu};
  for (int k = 0; k < 9; k++) {
    c[k].resize(n);
    if (d2h(c[k].data(), src[k], n * 8)) return fail(SGPU_ERR_HIP, "write_checkpoint: copy");
  }
  std::vector<int32_t> ind(n), type(n), fate(n);
  std::vector<uint32_t> draw(n);
  std::vector<uint64_t> id(n);
  if (d2h(ind.data(), pk_.ind, n * 4) || d2h(type.data(), pk_.type, n * 4) || d2h(fate.data(), pk_.fate, n * 4) ||
      d2h(draw.data(), pk_.draw, n * 4) || d2h(id.data(), pk_.id, n * 8))
    return fail(SGPU_ERR_HIP, "write_checkpoint: copy");
  std::vector<double> x(3 * n), D(3 * n);
  for (size_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) { x[3 * i + k] = c[k][i]; D[3 * i + k] = c[3 + k][i]; }
  const std::string pd = dir + "/particles/", sd = dir + "/state/";
  int bad = 0;
  bad |= write_npy(pd + "x.npy", "<f8", {n, 3}, x.data(), 24 * n);
  bad |= write_npy(pd + "D.npy", "<f8", {n, 3}, D.data(), 24 * n);
  bad |= write_npy(pd + "t.npy", "<f8", {n}, c[6].data(), 8 * n);
  bad |= write_npy(pd + "e.npy", "<f8", {n}, c[7].data(), 8 * n);
  bad |= write_npy(pd + "nu.npy", "<f8", {n}, c[8].data(), 8 * n);
  bad |= write_npy(pd + "type.npy", "<i4", {n}, type.data(), 4 * n);
  bad |= write_npy(pd + "ind.npy", "<i4", {n}, ind.data(), 4 * n);
  bad |= write_npy(pd + "fate.npy", "<i4", {n}, fate.data(), 4 * n);
  bad |= write_npy(pd + "draw.npy", "<u4", {n}, draw.data(), 4 * n);
  bad |= write_npy(pd + "id.npy", "<u8", {n}, id.data(), 8 * n);
  const size_t nz = nz_, nzn = (size_t)nz_ * nn_;
  const struct { const char* name; const double* d; size_t len; } st[] = {
      {"rho", d_rho_, nz}, {"T", zs_.T_gas, nz}, {"vol", d_vol_, nz}, {"r_out", d_r_out_, nz},
      {"e_abs", ta_.e_abs, nz}, {"J_nu", ta_.J_nu, nzn}, {"fx_rad", ta_.fx_rad, nz}, {"fy_rad", ta_.fy_rad, nz},
      {"fz_rad", ta_.fz_rad, nz}, {"fr_rad", ta_.fr_rad, nz}, {"L_radio_dep", ta_.L_radio_dep, nz},
      {"L_radio_emit", ta_.L_radio_emit, nz}, {"e_rad", ta_.e_rad, nz}};
  for (const auto& f : st) {
    std::vector<double> h = download(f.d, f.len);
    bad |= write_npy(sd + f.name + ".npy", "<f8", {f.len}, h.data(), 8 * f.len);
  }
  // the escape spectra accumulate over the steps of a time-dependent run
  // (transport::writeCheckpointSpectra, transport_output.cpp:492; read back at
  // transport_init.cpp:164)
  const struct { const char* name; const double* d; } sp[] = {
      {"opt_flux", ta_.optical.flux}, {"opt_click", ta_.optical.click},
      {"gam_flux", ta_.gamma.flux}, {"gam_click", ta_.gamma.click}};
  for (const auto& f : sp) {
    std::vector<double> h = download(f.d, n_spec_);
    bad |= write_npy(sd + f.name + ".npy", "<f8", {n_spec_}, h.data(), 8 * n_spec_);
  }
  FILE* m = fopen((dir + "/meta.json").c_str(), "w");
  if (!m) bad = 1;
  else {
    fprintf(m, "{\n \"format\": \"sedona-mi355x checkpoint 1\",\n \"config\": \"host\",\n \"seed\": %llu,\n"
               " \"n_zones\": %d,\n \"n_nu\": %d,\n \"n_packets\": %lld,\n \"t_now\": %.17g,\n"
               " \"grid_t_now\": %.17g,\n \"grid_min\": [%.17g, %.17g, %.17g],\n \"first_step\": %d,\n"
               " \"rng_seed\": %llu,\n \"step_index\": %llu\n}\n",
            (unsigned long long)p_.transport_rng_seed, nz_, nn_, (long long)n_particles_, t_now_, grid_.t_now,
            grid_.r_out.min, grid_.x_out.min, grid_.z_out.min, first_step_,
            (unsigned long long)p_.transport_rng_seed, (unsigned long long)step_index_);
    bad |= fclose(m) != 0;
  }
  return bad ? fail(SGPU_ERR_ARG, "write_checkpoint: cannot write the files") : SGPU_OK;
}

int transport::read_checkpoint(const std::string& dir)
{
  std::string js;
  if (FILE* m = fopen((dir + "/meta.json").c_str(), "r")) {
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), m)) > 0) js.append(buf, k);
    fclose(m);
  }
  if (js.find("\"sedona-mi355x checkpoint 1\"") == std::string::npos)
    return fail(SGPU_ERR_ARG, "read_checkpoint: no checkpoint of this format");
  if ((int)meta_number(js, "n_zones", -1) != nz_ || (int)meta_number(js, "n_nu", -1) != nn_)
    return fail(SGPU_ERR_ARG, "read_checkpoint: zone or frequency count differs");
  const int64_t n = (int64_t)meta_number(js, "n_packets", -1);
  if (n < 0 || n > pk_.capacity) return fail(SGPU_ERR_SIZE, "read_checkpoint: packet count exceeds capacity");
  std::vector<size_t> shape;
  std::vector<char> raw;
  auto load = [&](const std::string& path, const char* descr, size_t want, void* d, size_t item) {
    if (read_npy(path, descr, &shape, &raw) || raw.size() != want * item) return -1;
    return want ? sgpu_memcpy_h2d(ctx_, d, raw.data(), raw.size()) : 0;
  };
  const std::string pd = dir + "/particles/", sd = dir + "/state/";
  int bad = 0;
  if (n > 0) {
    double* dst[2][3] = {{pk_.x, pk_.y, pk_.z}, {pk_.Dx, pk_.Dy, pk_.Dz}};
    const char* nm[2] = {"x.npy", "D.npy"};
    for (int v = 0; v < 2; v++) {
      if (read_npy(pd + nm[v], "<f8", &shape, &raw) || raw.size() != (size_t)n * 24) { bad = 1; break; }
      const double* a = (const double*)raw.data();
      std::vector<double> col(n);
      for (int k = 0; k < 3; k++) {
        for (int64_t i = 0; i < n; i++) col[i] = a[3 * i + k];
        bad |= sgpu_memcpy_h2d(ctx_, dst[v][k], col.data(), 8 * n) != 0;
      }
    }
    bad |= load(pd + "t.npy", "<f8", n, pk_.t, 8) != 0;
    bad |= load(pd + "e.npy", "<f8", n, pk_.e, 8) != 0;
    bad |= load(pd + "nu.npy", "<f8", n, pk_.nu, 8) != 0;
    bad |= load(pd + "type.npy", "<i4", n, pk_.type, 4) != 0;
    bad |= load(pd + "ind.npy", "<i4", n, pk_.ind, 4) != 0;
    bad |= load(pd + "fate.npy", "<i4", n, pk_.fate, 4) != 0;
    bad |= load(pd + "draw.npy", "<u4", n, pk_.draw, 4) != 0;
    bad |= load(pd + "id.npy", "<u8", n, pk_.id, 8) != 0;
  }
  const size_t nz = nz_, nzn = (size_t)nz_ * nn_;
  const struct { const char* name; double* d; size_t len; } st[] = {
      {"rho", d_rho_, nz}, {"T", (double*)zs_.T_gas, nz}, {"vol", d_vol_, nz}, {"r_out", d_r_out_, nz},
      {"e_abs", ta_.e_abs, nz}, {"J_nu", ta_.J_nu, nzn}, {"fx_rad", ta_.fx_rad, nz}, {"fy_rad", ta_.fy_rad, nz},
      {"fz_rad", ta_.fz_rad, nz}, {"fr_rad", ta_.fr_rad, nz}, {"L_radio_dep", ta_.L_radio_dep, nz},
      {"L_radio_emit", ta_.L_radio_emit, nz}, {"e_rad", ta_.e_rad, nz}};
  for (const auto& f : st) bad |= load(sd + f.name + ".npy", "<f8", f.len, f.d, 8) != 0;
  const struct { const char* name; double* d; } sp[] = {
      {"opt_flux", ta_.optical.flux}, {"opt_click", ta_.optical.click},
      {"gam_flux", ta_.gamma.flux}, {"gam_click", ta_.gamma.click}};
  for (const auto& f : sp) bad |= load(sd + f.name + ".npy", "<f8", n_spec_, f.d, 8) != 0;
  if (bad) return fail(SGPU_ERR_ARG, "read_checkpoint: missing or malformed dataset");
  n_particles_ = n;
  t_now_ = meta_number(js, "t_now", t_now_);
  grid_.t_now = meta_number(js, "grid_t_now", grid_.t_now);
  grid_.r_out.min = meta_number(js, "grid_min", grid_.r_out.min);
  first_step_ = (int)meta_number(js, "first_step", first_step_);
  step_index_ = (uint64_t)meta_number(js, "step_index", (double)step_index_);
  return SGPU_OK;
}

int transport::write_plotfile(int iw, double tt, int write_mass_fracs, const std::string& dir)
{
  char name[64];
  snprintf(name, sizeof(name), "plt_%05d.dat", iw);
  // homologous zones: v = r_out / t (the velocity a homologous model file holds)
  std::vector<double> v(nz_);
  for (int i = 0; i < nz_; ++i) v[i] = g_.r_out[i] / grid_.t_now;
  int rc = write_plotfile_1d(dir + "/" + name, tt, g_.r_min, g_.r_out, download(d_rho_, nz_), v, T_gas(),
                             e_rad(), n_elec(),
                             L_radio_dep(), download(ta_.L_radio_emit, nz_),
                             write_mass_fracs ? &g_.X : nullptr, g_.elems_Z, g_.elems_A);
  return rc ? fail(SGPU_ERR_ARG, "write_plotfile: cannot write the zone file") : SGPU_OK;
}

}  // namespace sedona_gpu
