"""Text output in the reference's formats.

* ``write_spectrum_dat``: spectrum_array::print's ``<name>.dat``
  (src/transport/spectrum_array.cpp:253-311): a ``# n_t n_nu n_mu n_phi``
  header, then one line per (mu, phi, t, nu) bin, mu outermost, nu innermost,
  with a centre column for every axis of more than one bin, the flux divided
  by dnu dt / (n_mu n_phi) (a factor only for axes of more than one bin) and
  the click count.  transport::output_spectrum (transport_output.cpp:17-40)
  names the file ``<spectrum_name>_<it>`` (``_final`` for it < 0).
* ``write_plotfile_1d``: grid_1D_sphere::write_plotfile's ``plt_<iw>.dat``
  (src/grid/grid_1D_sphere.cpp:598-636): r_out, rho, v, T_gas,
  T_rad = (e_rad / a)^(1/4), n_elec, L_dep_nuc, L_emit_nuc and optionally the
  mass fractions, one zone per line.

A grid is the package's locate dict (``n``, ``min``, ``x`` = right edges);
centres and widths follow locate_array::center / delta
(src/utils/locate_array.h:53-57).
"""
import numpy as np

A_RAD = 7.5657e-15   # radiation constant, physical_constants.h


def _centers(g):
    x = np.asarray(g["x"], float)
    left = np.concatenate([[g["min"]], x[:-1]])
    return 0.5 * (left + x), x - left


def spectrum_file_name(base, it):
    """transport::output_spectrum's name: ``base_<it>``, ``base_final`` for it < 0."""
    return "%s_final" % base if it < 0 else "%s_%d" % (base, it)


def write_spectrum_dat(name, grids, flux, click, n_procs=1):
    """Write ``name + '.dat'``; grids = dict(time=, nu=, mu=, phi=)."""
    nt, nw, nm, npf = (int(grids[k]["n"]) for k in ("time", "nu", "mu", "phi"))
    flux = np.asarray(flux, float).reshape(nt, nw, nm, npf)
    click = np.asarray(click, float).reshape(nt, nw, nm, npf)
    tc, dt = _centers(grids["time"])
    wc, dw = _centers(grids["nu"])
    mc, _ = _centers(grids["mu"])
    pc, _ = _centers(grids["phi"])
    lines = ["# %d %d %d %d\n" % (nt, nw, nm, npf)]
    for k in range(nm):
        for m in range(npf):
            for i in range(nt):
                for j in range(nw):
                    s = ""
                    if nt > 1:
                        s += "%12.4e " % tc[i]
                    if nw > 1:
                        s += "%12.4e " % wc[j]
                    if nm > 1:
                        s += "%12.4f " % mc[k]
                    if npf > 1:
                        s += "%12.4f " % pc[m]
                    norm = 1.0 / (nm * npf)
                    if nw > 1:
                        norm *= dw[j]
                    if nt > 1:
                        norm *= dt[i]
                    s += "%12.5e %12.5e\n" % (flux[i, j, k, m] / norm, click[i, j, k, m] * n_procs)
                    lines.append(s)
    with open(name + ".dat", "w") as f:
        f.writelines(lines)


def write_plotfile_1d(path, t, r_min, r_out, rho, v, T_gas, e_rad, n_elec, L_dep, L_emit,
                      X=None, elems_Z=None, elems_A=None):
    """Write the 1D zone file; X [nz][n_elems] adds the mass-fraction columns."""
    nz = len(r_out)
    head = "#  %-12.12s %-15.15s %-15.15s %-15.15s %-15.15s %-15.15s %-15.15s %-15.15s" % (
        "r", "rho", "v", "T_gas", "T_rad", "n_elec", "L_dep_nuc", "L_emit_nuc")
    if X is not None:
        for z, a in zip(elems_Z, elems_A):
            head += " %-15.15s" % ("%d.%d" % (z, a))
    out = ["# t = %8.4e ; rmin = %8.4e\n" % (t, r_min), head + "\n"]
    for i in range(nz):
        T_rad = (e_rad[i] / A_RAD) ** 0.25
        s = "%12.8e  %12.8e  %12.8e  %12.8e  %12.8e  %12.8e  %12.8e  %12.8e" % (
            r_out[i], rho[i], v[i], T_gas[i], T_rad, n_elec[i], L_dep[i], L_emit[i])
        if X is not None:
            for x in X[i]:
                s += "  %12.8e" % x
        out.append(s + "\n")
    with open(path, "w") as f:
        f.writelines(out)


def output_spectrum(P, it, spectrum_name="spectrum", gamma_name=""):
    """transport::output_spectrum for a runtime.Problem: its optical (and, when
    gamma_name is set, gamma) spectrum as ``<name>_<it>.dat``."""
    for base, key in ((spectrum_name, "opt"), (gamma_name, "gam")):
        if base:
            write_spectrum_dat(spectrum_file_name(base, it), P.spec_grids, P.get(key + "_flux"),
                               P.get(key + "_click"))


def read_spectrum_dat(path):
    """Parse a spectrum ``.dat``: (header counts, float array of the rows)."""
    with open(path) as f:
        head = f.readline().split()
        counts = tuple(int(v) for v in head[1:5])
        rows = np.loadtxt(f, ndmin=2)
    return counts, rows
