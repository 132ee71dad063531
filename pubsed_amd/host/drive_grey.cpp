// drive_grey.cpp -- a host C++ caller of the drop-in transport class: the
// grey 1D 56Ni ejecta (100 shells to 2e9 cm/s at 20 d, rho ~ exp(-v/3e8)
// normalised to 1.4 Msun, 56Ni inside 0.6 Msun, grey kappa 0.1 + Thomson),
// one steady-state step, results printed as one JSON line; with an output
// directory, the spectra (spectrum_0.dat, gamma_0.dat) and the zone file
// plt_00000.dat are written there in the reference's text formats.
//
// With a checkpoint directory as the fourth argument it instead runs three
// time-dependent steps of dt = 2 d twice -- straight through, and with a
// checkpoint written after the first step and read into a fresh transport
// that runs the other two -- and prints both runs' results, including the
// escape spectrum (half-day time bins, 4 mu bins) accumulated over the steps.
//
// usage: drive_grey [n_packets] [precision 0|1] [output_dir] [checkpoint_dir]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sedona_transport.hpp"

int main(int argc, char** argv)
{
  using namespace sedona_gpu;
  const int64_t n_packets = argc > 1 ? atoll(argv[1]) : 100000;
  const int precision = argc > 2 ? atoi(argv[2]) : SGPU_PREC_MIXED;
  const std::string out_dir = argc > 3 ? argv[3] : "";
  const std::string chk_dir = argc > 4 ? argv[4] : "";
  const double PI_REF = 3.14159, M_SUN = 1.98892e33, DAY = 86400.0;
  const double kappa_T = 0.66523e-24 * 0.5 / 1.67262158e-24;
  const int nz = 100;
  const double v_max = 2e9, t = 20 * DAY, v_e = 3e8;

  Grid1D g;
  g.t_now = t;
  g.elems_Z = {28, 27, 26, 14};
  g.elems_A = {56, 56, 56, 28};
  std::vector<double> vc(nz), vol(nz);
  double mass = 0;
  for (int i = 0; i < nz; ++i) {
    double vo = v_max * (i + 1) / nz;
    vc[i] = vo - 0.5 * v_max / nz;
    g.r_out.push_back(vo * t);
    double r0 = i == 0 ? 0.0 : g.r_out[i - 1];
    vol[i] = 4.0 * PI_REF / 3.0 * (g.r_out[i] * g.r_out[i] * g.r_out[i] - r0 * r0 * r0);
    g.rho.push_back(std::exp(-vc[i] / v_e));
    mass += g.rho[i] * vol[i];
  }
  double menc = 0;
  for (int i = 0; i < nz; ++i) {
    g.rho[i] *= 1.4 * M_SUN / mass;
    double mc = menc + 0.5 * g.rho[i] * vol[i];
    menc += g.rho[i] * vol[i];
    bool ni = mc <= 0.6 * M_SUN;
    g.X.insert(g.X.end(), {ni ? 1.0 : 0.0, 0.0, 0.0, ni ? 0.0 : 1.0});
    g.T_gas.push_back(1.5e4 + (4e3 - 1.5e4) * vc[i] / v_max);
  }
  Params p;
  p.transport_nu_grid[0] = 1e14;
  p.transport_nu_grid[1] = 1e16;
  p.n_nu = 1024;
  p.opacity_grey_opacity = 0.1 + kappa_T;
  p.opacity_epsilon = 0.1 / (0.1 + kappa_T);
  p.particles_n_emit_radioactive = n_packets;
  p.particles_max_total = n_packets;
  p.precision = precision;
  p.spectrum_name = out_dir.empty() ? "" : out_dir + "/spectrum";
  p.gamma_name = out_dir.empty() ? "" : out_dir + "/gamma";

  if (!chk_dir.empty()) {
    p.transport_steady_iterate = 0;
    p.particles_max_total = 3 * n_packets;
    // the escape spectrum in half-day time bins and 4 mu bins, accumulating
    // over the steps
    p.spectrum_time_grid[0] = 18 * DAY;
    p.spectrum_time_grid[1] = 28 * DAY;
    p.spectrum_time_grid[2] = 0.5 * DAY;
    p.spectrum_n_mu = 4;
    const double dt = 2 * DAY;
    auto summary = [&](transport& tr) {
      double dep = 0, eabs = 0, J = 0, spec = 0, spec_w = 0, clicks = 0;
      auto L = tr.L_radio_dep();
      auto A = tr.e_abs();
      for (int i = 0; i < nz; ++i) { dep += L[i] * vol[i]; eabs += A[i] * vol[i]; }
      for (double j : tr.J_nu()) J += j;
      auto S = tr.optical_spectrum();
      auto C = tr.optical_clicks();
      for (size_t i = 0; i < S.size(); ++i) { spec += S[i]; spec_w += S[i] * (double)(i + 1); clicks += C[i]; }
      char b[768];
      snprintf(b, sizeof(b), "{\"alive\": %lld, \"escaped\": %lld, \"t_now\": %.17g, \"L_dep\": %.17g, "
               "\"E_abs_rate\": %.17g, \"J_sum\": %.17g, \"spec_flux\": %.17g, \"spec_flux_weighted\": %.17g, "
               "\"spec_clicks\": %.17g, \"spec_bins\": %zu}", (long long)tr.n_particles(),
               (long long)tr.n_escaped_last(), tr.t_now(), dep, eabs, J, spec, spec_w, clicks, S.size());
      return std::string(b);
    };
    transport a, b, c;
    if (a.init(p, g, nullptr) || b.init(p, g, nullptr) || c.init(p, g, nullptr)) {
      fprintf(stderr, "init: %s%s%s\n", a.error().c_str(), b.error().c_str(), c.error().c_str());
      return 1;
    }
    for (int s = 0; s < 3; ++s)
      if (a.step(dt) != SGPU_OK) { fprintf(stderr, "step: %s\n", a.error().c_str()); return 1; }
    if (b.step(dt) != SGPU_OK || b.write_checkpoint(chk_dir) != SGPU_OK) {
      fprintf(stderr, "checkpoint: %s\n", b.error().c_str());
      return 1;
    }
    if (c.read_checkpoint(chk_dir) != SGPU_OK) { fprintf(stderr, "restart: %s\n", c.error().c_str()); return 1; }
    for (int s = 0; s < 2; ++s)
      if (c.step(dt) != SGPU_OK) { fprintf(stderr, "step: %s\n", c.error().c_str()); return 1; }
    printf("{\"straight\": %s, \"restarted\": %s}\n", summary(a).c_str(), summary(c).c_str());
    return 0;
  }

  transport tr;
  if (tr.init(p, g, nullptr) != SGPU_OK) { fprintf(stderr, "init: %s\n", tr.error().c_str()); return 1; }
  if (tr.step(1.0) != SGPU_OK) { fprintf(stderr, "step: %s\n", tr.error().c_str()); return 1; }
  if (!out_dir.empty() && (tr.output_spectrum(0) != SGPU_OK || tr.write_plotfile(0, t, 1, out_dir) != SGPU_OK)) {
    fprintf(stderr, "output: %s\n", tr.error().c_str());
    return 1;
  }
  double dep = 0, erad = 0, eabs = 0, spec = 0;
  auto L = tr.L_radio_dep();
  auto E = tr.e_rad();
  auto A = tr.e_abs();
  auto S = tr.optical_spectrum();
  for (int i = 0; i < nz; ++i) { dep += L[i] * vol[i]; erad += E[i] * vol[i]; eabs += A[i] * vol[i]; }
  for (double s : S) spec += s;
  printf("{\"packets\": %lld, \"escaped\": %lld, \"alive\": %lld, \"L_dep\": %.17g, \"E_rad\": %.17g, "
         "\"E_abs_rate\": %.17g, \"optical_flux\": %.17g}\n",
         (long long)n_packets, (long long)tr.n_escaped_last(), (long long)tr.n_particles(), dep, erad, eabs,
         spec);
  return 0;
}
