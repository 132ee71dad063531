// sedona_transport.hpp -- host C++ drop-in for Sedona's transport step on
// one MI355X.  The class keeps the reference's member names and step order
// (transport::step, src/transport/transport.cpp:26-140): set_opacity,
// wipe_radiation, emit_particles, the propagation loop, clean_up of escaped
// and absorbed packets, reduce_radiation; each is one C-ABI call
// (include/sedona_gpu.h) over all zones or all packets.  Device buffers are
// owned by the object; the caller describes the grid and gas on the host.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sedona_gpu.h"

namespace sedona_gpu {

// parameters, named after sedona_defaults.lua
struct Params {
  double transport_nu_grid[3] = {1e14, 1e16, 0};   // {nu_1, nu_2, del} log grid
  int n_nu = 1024;                                  // bins (del chosen to give exactly n_nu)
  int transport_steady_iterate = 1;
  int transport_radiative_equilibrium = 0;
  int transport_fix_Tgas_during_transport = 0;
  int transport_set_Tgas_to_Trad = 0;
  int transport_store_Jnu = 1;
  int transport_boundary_in_reflect = 0;
  int transport_boundary_out_reflect = 0;
  double transport_fleck_alpha = 0;
  int transport_use_ddmc = 0;                 // 0 off, 2 DDMC (1D, FP64 form)
  double transport_ddmc_tau_threshold = 100;
  double opacity_grey_opacity = 0;
  double opacity_epsilon = 1.0;
  int opacity_electron_scattering = 0;
  int opacity_fuzz_expansion = 0;
  int opacity_line_expansion = 0;
  int opacity_free_free = 0;
  int opacity_bound_free = 0;
  int opacity_no_ground_recomb = 0;
  int opacity_bound_bound = 0;                // Voigt-profile line opacity (level-resolved lines)
  int opacity_use_nlte = 0;                   // NLTE gas state from the last step's J_nu
  std::vector<int32_t> opacity_atoms_in_nlte; // per-atom flags (element order)
  int opacity_use_collisions_nlte = 1;
  double line_velocity_width = 0;
  double opacity_minimum_extinction = 0;     // bound-bound lines below this alpha_0/(nu_0^2 dnu) are skipped
  double opacity_maximum_opacity = 1e40;
  double limits_temp_min = 1000;
  double limits_temp_max = 1e8;
  int64_t particles_n_initialize = 0;        // initial radiation field packets (initialize_particles)
  int64_t particles_n_emit_radioactive = 0;
  // inner boundary source (setup_core_emission / emit_inner_source)
  int64_t core_n_emit = 0;
  double core_radius = 0;                      // also the grid's inner absorbing boundary
  double core_temperature = 0;
  double core_luminosity = 0;
  double core_photon_frequency = 0;
  int64_t particles_max_total = 10000000;
  uint64_t transport_rng_seed = 20261203;
  int precision = SGPU_PREC_MIXED;
  double spectrum_time_grid[3] = {0, 1, 1};  // {start, stop, del}; start >= stop: one catch-all bin
  int spectrum_n_mu = 1;                      // mu = D_z bins on [-1, 1]
  int spectrum_n_phi = 1;                     // phi bins on [0, 2 pi]
  std::string spectrum_name = "spectrum";   // output file stem; "" writes nothing
  std::string gamma_name = "";
  int spectrum_suppress_txt = 0;
};

// a locate_array on the host: right edges x and the left edge of bin 0
struct HostGrid {
  double min = 0;
  std::vector<double> x;
  int size() const { return (int)x.size(); }
  double center(int i) const { return 0.5 * ((i == 0 ? min : x[i - 1]) + x[i]); }
  double delta(int i) const { return x[i] - (i == 0 ? min : x[i - 1]); }
};

// spectrum_array::print's text file <name>.dat (spectrum_array.cpp:253-311);
// flux / click are [n_t][n_nu][n_mu][n_phi].  Returns 0, or -1 when the file
// cannot be written.
int write_spectrum_dat(const std::string& name, const HostGrid& time, const HostGrid& nu,
                       const HostGrid& mu, const HostGrid& phi, const std::vector<double>& flux,
                       const std::vector<double>& click, int n_procs = 1);

// grid_1D_sphere::write_plotfile's text file (grid_1D_sphere.cpp:598-636)
int write_plotfile_1d(const std::string& path, double t, double r_min, const std::vector<double>& r_out,
                      const std::vector<double>& rho, const std::vector<double>& v,
                      const std::vector<double>& T_gas, const std::vector<double>& e_rad,
                      const std::vector<double>& n_elec, const std::vector<double>& L_dep,
                      const std::vector<double>& L_emit, const std::vector<double>* X,
                      const std::vector<int32_t>& elems_Z, const std::vector<int32_t>& elems_A);

// NumPy .npy (format 1.0) datasets of a checkpoint; descr such as "<f8",
// "<i4", "<u4", "<u8".  read_npy checks descr and C order and returns the
// shape and the raw data.
int write_npy(const std::string& path, const char* descr, const std::vector<size_t>& shape, const void* data,
              size_t bytes);
int read_npy(const std::string& path, const char* descr, std::vector<size_t>* shape, std::vector<char>* data);

// 1D spherical homologous grid, host side (grid_1D_sphere's fields)
struct Grid1D {
  std::vector<double> r_out;      // shell outer radii
  double r_min = 0;
  double t_now = 0;               // grid clock (velocity = r / t_now)
  std::vector<double> rho, T_gas;
  std::vector<double> e_rad;      // radiation energy density; empty: a T_gas^4 (grid_1D_sphere.cpp:354)
  std::vector<double> X;          // [nz][n_elems]
  std::vector<int32_t> elems_Z, elems_A;
  int n_zones() const { return (int)r_out.size(); }
};

// atomic data in the C-ABI layout (see sgpu_atoms); host vectors
struct AtomicData {
  std::vector<int32_t> lev_off, ion_off, lev_g, lev_ion, atom_zero_eps, line_ion, line_bin;
  std::vector<double> lev_E, ion_chi, line_nu, line_gf, line_El;
  std::vector<int64_t> line_bin_off;
  // level-resolved lines (AtomicData.h:12-56 lines_): lower / upper level,
  // frequency, f_lu, A_ul, B_ul, B_lu and their bin CSR
  std::vector<int32_t> bline_l, bline_u, bline_bin;
  std::vector<double> bline_nu, bline_f, bline_A, bline_Bul, bline_Blu;
  std::vector<int64_t> bline_bin_off;
  int max_levels = 0;
};

class transport {
 public:
  transport() = default;
  ~transport();
  transport(const transport&) = delete;
  transport& operator=(const transport&) = delete;

  // transport::init (src/transport/transport_init.cpp:44-427), restricted to
  // the hot path: frequency grid, tables, tallies, packet buffers
  int init(const Params& p, const Grid1D& g, const AtomicData* atoms, int device = 0);
  // transport::step (transport.cpp:26-140)
  int step(double dt);

  // results (copied to the host)
  std::vector<double> J_nu() const;             // [nz][n_nu], normalised
  std::vector<double> e_rad() const;
  std::vector<double> e_abs() const;
  std::vector<double> L_radio_dep() const;
  std::vector<double> optical_spectrum() const; // flux per [t][nu][mu][phi] bin (raw sums of energy)
  std::vector<double> optical_clicks() const;   // packets per bin
  std::vector<double> abs_opacity() const;
  std::vector<double> T_gas() const;
  std::vector<double> n_elec() const;
  // transport::output_spectrum (transport_output.cpp:17-40): the optical and
  // gamma spectra as <name>_<it>.dat (<name>_final.dat for it < 0); the
  // HDF5 twin is not written
  int output_spectrum(int it);
  // grid_1D_sphere::write_plotfile's plt_<iw>.dat, under dir
  int write_plotfile(int iw, double tt, int write_mass_fracs = 0, const std::string& dir = ".");
  // transport::writeCheckpointParticles / readCheckpointParticles
  // (transport_output.cpp:300-365, 501-590) with .npy datasets under a
  // directory in place of HDF5 groups: particles/{x,D (n x 3),type,ind,t,e,
  // nu,fate,draw,id}.npy (draw and id are the packet's RNG stream state),
  // state/{rho,T,vol,r_out and the radiation tallies}.npy and meta.json with
  // the clocks -- the layout pubsed_amd.checkpoint reads and writes.  A
  // transport initialised with the same parameters and grid and restored
  // from a checkpoint continues the run packet for packet.
  int write_checkpoint(const std::string& dir);
  int read_checkpoint(const std::string& dir);
  int64_t n_particles() const { return n_particles_; }
  int64_t n_escaped_last() const { return n_escaped_; }
  const std::string& error() const { return err_; }
  double t_now() const { return t_now_; }

 private:
  int set_opacity(double dt);
  int wipe_radiation();
  int emit_particles(double dt);
  int propagate_all(double dt);
  int clean_up_particle_vector();
  int reduce_radiation(double dt);
  int fail(int rc, const char* what);
  void* dalloc(size_t bytes);
  template <class T> T* upload(const std::vector<T>& v);
  std::vector<double> download(const double* d, size_t n) const;

  sgpu_ctx* ctx_ = nullptr;
  Params p_;
  int nz_ = 0, nn_ = 0, ne_ = 0;
  double t_now_ = 0;
  int first_step_ = 1;
  uint64_t step_index_ = 0;
  int64_t n_particles_ = 0, n_escaped_ = 0;
  std::string err_;
  std::vector<void*> owned_;
  std::vector<double> nu_edges_;
  HostGrid spec_t_, spec_mu_, spec_phi_;   // spectrum_array's time / mu / phi grids
  size_t n_spec_ = 0;                      // n_t n_nu n_mu n_phi
  Grid1D g_;
  sgpu_locate nu_{};
  sgpu_grid grid_{};
  sgpu_zone_state zs_{};
  sgpu_atoms at_{};
  bool have_atoms_ = false;
  sgpu_opacity_params op_{};
  sgpu_opacity_out out_{};
  sgpu_tables tab_{};
  sgpu_tallies ta_{};
  sgpu_transport_params tp_{};
  sgpu_packets pk_{}, scr_{};
  double* d_rho_ = nullptr;
  double* d_r_out_ = nullptr;
  double* d_vol_ = nullptr;
};

}  // namespace sedona_gpu
