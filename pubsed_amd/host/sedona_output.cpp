// sedona_output.cpp -- the reference's text output formats, written on the
// host from tallies copied back from the device.
//   write_spectrum_dat: spectrum_array::print (src/transport/spectrum_array.cpp:253-311)
//   write_plotfile_1d:  grid_1D_sphere::write_plotfile (src/grid/grid_1D_sphere.cpp:598-636)
// and the .npy datasets of a checkpoint (write_npy / read_npy, NumPy format
// version 1.0: magic, header length, a Python-literal header dict, raw data).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "sedona_transport.hpp"

namespace sedona_gpu {

namespace {
constexpr double A_RAD = 7.5657e-15;   // radiation constant, physical_constants.h
}

int write_spectrum_dat(const std::string& name, const HostGrid& time, const HostGrid& nu,
                       const HostGrid& mu, const HostGrid& phi, const std::vector<double>& flux,
                       const std::vector<double>& click, int n_procs)
{
  const int nt = time.size(), nw = nu.size(), nm = mu.size(), np = phi.size();
  const size_t n = (size_t)nt * nw * nm * np;
  if (flux.size() != n || click.size() != n) return -1;
  FILE* out = fopen((name + ".dat").c_str(), "w");
  if (!out) return -1;
  fprintf(out, "# %d %d %d %d\n", nt, nw, nm, np);
  // mu outermost, then phi, time, frequency; index = ((t nw + l) nm + m) np + p
  for (int k = 0; k < nm; k++)
    for (int m = 0; m < np; m++)
      for (int i = 0; i < nt; i++)
        for (int j = 0; j < nw; j++) {
          const size_t id = (((size_t)i * nw + j) * nm + k) * np + m;
          if (nt > 1) fprintf(out, "%12.4e ", time.center(i));
          if (nw > 1) fprintf(out, "%12.4e ", nu.center(j));
          if (nm > 1) fprintf(out, "%12.4f ", mu.center(k));
          if (np > 1) fprintf(out, "%12.4f ", phi.center(m));
          double norm = 1.0 / (nm * np);
          if (nw > 1) norm *= nu.delta(j);
          if (nt > 1) norm *= time.delta(i);
          fprintf(out, "%12.5e %12.5e\n", flux[id] / norm, click[id] * n_procs);
        }
  return fclose(out) == 0 ? 0 : -1;
}

int write_plotfile_1d(const std::string& path, double t, double r_min, const std::vector<double>& r_out,
                      const std::vector<double>& rho, const std::vector<double>& v,
                      const std::vector<double>& T_gas, const std::vector<double>& e_rad,
                      const std::vector<double>& n_elec, const std::vector<double>& L_dep,
                      const std::vector<double>& L_emit, const std::vector<double>* X,
                      const std::vector<int32_t>& elems_Z, const std::vector<int32_t>& elems_A)
{
  FILE* out = fopen(path.c_str(), "w");
  if (!out) return -1;
  const size_t nz = r_out.size(), ne = elems_Z.size();
  fprintf(out, "# t = %8.4e ; rmin = %8.4e\n", t, r_min);
  fprintf(out, "#  %-12.12s %-15.15s %-15.15s %-15.15s %-15.15s %-15.15s %-15.15s %-15.15s", "r", "rho", "v",
          "T_gas", "T_rad", "n_elec", "L_dep_nuc", "L_emit_nuc");
  if (X) {
    for (size_t j = 0; j < ne; j++) {
      char id[32];
      snprintf(id, sizeof(id), "%d.%d", elems_Z[j], elems_A[j]);
      fprintf(out, " %-15.15s", id);
    }
  }
  fprintf(out, "\n");
  for (size_t i = 0; i < nz; i++) {
    const double T_rad = pow(e_rad[i] / A_RAD, 0.25);
    fprintf(out, "%12.8e  %12.8e  %12.8e  %12.8e  %12.8e  %12.8e  %12.8e  %12.8e", r_out[i], rho[i], v[i],
            T_gas[i], T_rad, n_elec[i], L_dep[i], L_emit[i]);
    if (X)
      for (size_t j = 0; j < ne; j++) fprintf(out, "  %12.8e", (*X)[i * ne + j]);
    fprintf(out, "\n");
  }
  return fclose(out) == 0 ? 0 : -1;
}

int write_npy(const std::string& path, const char* descr, const std::vector<size_t>& shape, const void* data,
              size_t bytes)
{
  std::string sh = "(";
  for (size_t i = 0; i < shape.size(); i++) sh += std::to_string(shape[i]) + (shape.size() == 1 ? "," : i + 1 < shape.size() ? ", " : "");
  sh += ")";
  std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
  // magic (6) + version (2) + length (2) + header, padded with spaces and a
  // newline to a multiple of 64 bytes
  size_t total = 10 + hdr.size() + 1;
  hdr.append((64 - total % 64) % 64, ' ');
  hdr += '\n';
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return -1;
  const unsigned char head[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  const uint16_t len = (uint16_t)hdr.size();
  const unsigned char l2[2] = {(unsigned char)(len & 0xff), (unsigned char)(len >> 8)};
  bool ok = fwrite(head, 1, 8, f) == 8 && fwrite(l2, 1, 2, f) == 2 && fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() &&
            (bytes == 0 || fwrite(data, 1, bytes, f) == bytes);
  return (fclose(f) == 0 && ok) ? 0 : -1;
}

int read_npy(const std::string& path, const char* descr, std::vector<size_t>* shape, std::vector<char>* data)
{
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return -1;
  unsigned char head[10];
  int rc = -1;
  if (fread(head, 1, 10, f) == 10 && head[0] == 0x93 && memcmp(head + 1, "NUMPY", 5) == 0 && head[6] == 1) {
    const size_t len = head[8] | ((size_t)head[9] << 8);
    std::string hdr(len, '\0');
    if (fread(&hdr[0], 1, len, f) == len) {
      const std::string want = std::string("'descr': '") + descr + "'";
      const size_t sp = hdr.find("'shape': (");
      if (hdr.find(want) != std::string::npos && hdr.find("'fortran_order': False") != std::string::npos &&
          sp != std::string::npos) {
        shape->clear();
        const char* c = hdr.c_str() + sp + 10;
        size_t count = 1;
        while (*c && *c != ')') {
          char* end;
          const unsigned long long v = strtoull(c, &end, 10);
          if (end == c) { ++c; continue; }
          shape->push_back((size_t)v);
          count *= (size_t)v;
          c = end;
        }
        size_t item = (size_t)atoi(descr + 2);
        data->resize(count * item);
        rc = (count == 0 || fread(data->data(), 1, data->size(), f) == data->size()) ? 0 : -1;
      }
    }
  }
  fclose(f);
  return rc;
}

}  // namespace sedona_gpu
