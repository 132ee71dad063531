"""The reference's text output formats (CPU only).

* spectrum_array::print's ``.dat`` (src/transport/spectrum_array.cpp:253-311):
  header counts, one centre column per axis of more than one bin, mu
  outermost and nu innermost (index = t a1 + l a2 + m a3 + p, :206-209), flux
  divided by dnu dt / (n_mu n_phi), clicks; the single-bin axes of the
  default spectrum (spectrum_time_grid {0,1,1}, one mu and phi bin) add
  neither a column nor a factor.
* grid_1D_sphere::write_plotfile's ``plt_*.dat`` (grid_1D_sphere.cpp:598-636).
* The host C++ writers (pubsed_amd/host/sedona_output.cpp) produce the same
  bytes as the Python ones for the same data.
"""
import os
import subprocess

import numpy as np
import pytest

from pubsed_amd import output as O, workloads as W

HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pubsed_amd", "host")


def grids(nt=3, nw=5, nm=2, npf=1):
    return dict(time=W.lin_init_n(0.0, 3e6, nt) if nt > 1 else W.lin_init_del(0.0, 1.0, 1.0),
                nu=W.nu_log_grid(1e14, 1e16, nw), mu=W.lin_init_n(-1, 1, nm),
                phi=W.lin_init_n(0, 2 * W.PI_REF, npf))


def test_spectrum_dat_layout_and_normalisation(tmp_path):
    g = grids()
    nt, nw, nm, npf = 3, 5, 2, 1
    rng = np.random.default_rng(1)
    flux = rng.uniform(1, 2, nt * nw * nm * npf) * 1e40
    click = rng.integers(0, 1000, nt * nw * nm * npf).astype(float)
    name = str(tmp_path / "spectrum_0")
    O.write_spectrum_dat(name, g, flux, click)
    counts, rows = O.read_spectrum_dat(name + ".dat")
    assert counts == (nt, nw, nm, npf)
    assert rows.shape == (nt * nw * nm * npf, 3 + 2)    # t, nu, mu centres + flux, click
    F = flux.reshape(nt, nw, nm, npf)
    C = click.reshape(nt, nw, nm, npf)
    tx = np.asarray(g["time"]["x"]); tl = np.concatenate([[g["time"]["min"]], tx[:-1]])
    wx = np.asarray(g["nu"]["x"]); wl = np.concatenate([[g["nu"]["min"]], wx[:-1]])
    mx = np.asarray(g["mu"]["x"]); ml = np.concatenate([[g["mu"]["min"]], mx[:-1]])
    r = 0
    for k in range(nm):            # mu outermost
        for i in range(nt):
            for j in range(nw):    # nu innermost
                row = rows[r]
                assert abs(row[0] / (0.5 * (tl[i] + tx[i])) - 1) < 1e-4
                assert abs(row[1] / (0.5 * (wl[j] + wx[j])) - 1) < 1e-4
                assert abs(row[2] - 0.5 * (ml[k] + mx[k])) < 1e-4
                want = F[i, j, k, 0] / ((wx[j] - wl[j]) * (tx[i] - tl[i]) / (nm * npf))
                assert abs(row[3] / want - 1) < 1e-5
                assert row[4] == C[i, j, k, 0]
                r += 1


def test_default_spectrum_has_only_frequency_column(tmp_path):
    g = grids(nt=1, nw=8, nm=1, npf=1)
    flux = np.arange(1, 9) * 1e39
    name = str(tmp_path / "spectrum_final")
    O.write_spectrum_dat(name, g, flux, np.ones(8))
    counts, rows = O.read_spectrum_dat(name + ".dat")
    assert counts == (1, 8, 1, 1) and rows.shape == (8, 3)
    wx = np.asarray(g["nu"]["x"]); wl = np.concatenate([[g["nu"]["min"]], wx[:-1]])
    assert np.allclose(rows[:, 1], flux / (wx - wl), rtol=1e-5, atol=0)
    assert O.spectrum_file_name("spectrum", -1) == "spectrum_final"
    assert O.spectrum_file_name("spectrum", 7) == "spectrum_7"


def test_plotfile_columns(tmp_path):
    nz = 4
    r = np.linspace(1e14, 4e14, nz)
    e_rad = np.array([7.5657e-15 * 1e16, 7.5657e-15 * 16e16, 1.0, 0.0])
    X = np.array([[0.7, 0.3]] * nz)
    path = str(tmp_path / "plt_00000.dat")
    O.write_plotfile_1d(path, 1.728e6, 0.0, r, np.full(nz, 1e-13), r / 1.728e6, np.full(nz, 1e4), e_rad,
                        np.full(nz, 1e9), np.zeros(nz), np.ones(nz), X, [28, 14], [56, 28])
    lines = open(path).read().splitlines()
    assert lines[0] == "# t = 1.7280e+06 ; rmin = 0.0000e+00"
    assert lines[1].split()[1:] == ["r", "rho", "v", "T_gas", "T_rad", "n_elec", "L_dep_nuc", "L_emit_nuc",
                                    "28.56", "14.28"]
    d = np.loadtxt(path)
    assert d.shape == (nz, 10)
    assert np.allclose(d[:2, 4], [1e4, 2e4], rtol=1e-8)     # T_rad = (e_rad / a)^(1/4)
    assert np.allclose(d[:, 2], r / 1.728e6, rtol=1e-8)


CPP_MAIN = r'''
#include <cstdio>
#include <vector>
#include "sedona_transport.hpp"
using namespace sedona_gpu;
static std::vector<double> rd(const char* p, size_t n) {
  std::vector<double> v(n); FILE* f = fopen(p, "rb"); fread(v.data(), 8, n, f); fclose(f); return v; }
int main(int argc, char** argv) {
  std::vector<double> g = rd(argv[1], 4 + 4 + 6 + 3 + 2);
  HostGrid t{g[4], {g[5], g[6], g[7]}}, nu{g[8], {g[9], g[10], g[11], g[12], g[13]}};
  HostGrid mu{g[14], {g[15], g[16]}}, ph{g[17], {g[18]}};
  std::vector<double> flux = rd(argv[2], 30), click = rd(argv[3], 30);
  if (write_spectrum_dat(argv[4], t, nu, mu, ph, flux, click)) return 1;
  std::vector<double> r = rd(argv[5], 4 * 10);
  std::vector<double> c[10];
  for (int k = 0; k < 10; ++k) c[k].assign(r.begin() + 4 * k, r.begin() + 4 * k + 4);
  std::vector<double> X = rd(argv[6], 8);
  std::vector<int32_t> Z = {28, 14}, A = {56, 28};
  return write_plotfile_1d(argv[7], 1.728e6, 0.0, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], &X, Z, A);
}
'''


def test_host_cpp_writers_match_python_bytes(tmp_path):
    g = grids()
    gv = [3, 5, 2, 1, g["time"]["min"], *g["time"]["x"], g["nu"]["min"], *g["nu"]["x"],
          g["mu"]["min"], *g["mu"]["x"], g["phi"]["min"], *g["phi"]["x"]]
    rng = np.random.default_rng(2)
    flux = rng.uniform(0, 1, 30) * 1e41
    flux[3] = 0
    click = rng.integers(0, 10 ** 6, 30).astype(float)
    cols = rng.uniform(0.5, 2, (10, 4)) * np.array([1e15, 1e-13, 1e9, 1e4, 1e-2, 1e9, 1e40, 1e41, 1, 1])[:, None]
    X = rng.uniform(0, 1, 8)
    files = {}
    for k, v in (("g", gv), ("f", flux), ("c", click), ("r", cols.ravel()), ("x", X)):
        files[k] = str(tmp_path / (k + ".bin"))
        np.asarray(v, np.float64).tofile(files[k])
    src = tmp_path / "m.cpp"
    src.write_text(CPP_MAIN)
    exe = str(tmp_path / "m")
    b = subprocess.run(["g++", "-O1", "-std=c++17", "-I" + HOST, "-o", exe, str(src),
                        os.path.join(HOST, "sedona_output.cpp")], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([exe, files["g"], files["f"], files["c"], str(tmp_path / "cpp_spec"), files["r"],
                        files["x"], str(tmp_path / "cpp_plt.dat")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    O.write_spectrum_dat(str(tmp_path / "py_spec"), g, flux, click)
    c = cols
    O.write_plotfile_1d(str(tmp_path / "py_plt.dat"), 1.728e6, 0.0, c[0], c[1], c[2], c[3], c[4], c[5], c[6],
                        c[7], X.reshape(4, 2), [28, 14], [56, 28])
    assert (tmp_path / "cpp_spec.dat").read_bytes() == (tmp_path / "py_spec.dat").read_bytes()
    assert (tmp_path / "cpp_plt.dat").read_bytes() == (tmp_path / "py_plt.dat").read_bytes()
