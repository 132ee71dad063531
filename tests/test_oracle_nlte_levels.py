"""Pins of the NLTE-coupled gas state on the reference's own few-level test
problems (CPU only; the GPU twins are in test_gpu_nlte_levels.py).

Both problems are pure hydrogen at T = 5e4 K and rho = 1e-20 g cm^-3
(optically thin) in the diluted field J_nu = W(r) B_nu(T_s) of a T_s = 5e4 K
photosphere of radius r_p = 3e13 cm, W = (1 - sqrt(1 - (r_p/r)^2)) / 2,
solved through the full opacity build (GasState::solve_state under
opacity_use_nlte: radiative rates from J_nu, set_rates, the LU solve and the
n_e root), collisions off as in the analytic derivations.

* Three bound levels plus continuum (tests/3level_nlte: levels n = 1, 2, 3
  at 0, 10.20, 12.09 eV, g = 2, 8, 18, A = 4.696e8, 5.572e7, 4.408e7 s^-1,
  the grid {0.5e14, 2e16, 0.0005, log} and line width 1e8 cm/s of its
  param.lua).  (a) The closed form of tests/3level_nlte/problem_writeup.pdf
  eqs. (15)-(18) -- n1/n3 = (AD - CF)/(BD + CE), n1/n2 = (AD - CF)/(AE + BF)
  -- evaluated with the build's own photoionisation rates I_i, recombination
  coefficients alpha_i and line mean intensities J-bar, to 1e-12: every term
  of the rate matrix enters it, so a dropped or misplaced rate fails.  (b)
  The committed analytic table (NLTE_analytic_solution.dat, 128 radii) within
  the writeup's own account of the reference ("within 1% over most of the
  domain, rising above 5% at the innermost zones"): its ratios come from the
  writeup's own photoionisation cross sections, these from the hydrogenic
  nu^-3 table of AtomicSpecies (6.3e-18 n_eff cm^2): measured 2.0% / 3.7%
  (n1/n2, n1/n3) innermost, 0.34% / 0.58% mean, 0.23% / 0.37% beyond r = 2 r_p.
* Two bound levels (tests/2level_nlte: 0 and 10.2 eV, g = 2, 8,
  A = 4.7e8 s^-1, line width 5e7 cm/s, the 100 zone centres of its
  vacuum.mod) with a continuum 60 eV up, whose weak bound-free rates keep the
  system regular: (a) the two-level-plus-continuum closed form (writeup eq.
  (11) in terms of J-bar) to 1e-12; (b) the formula of
  tests/2level_nlte/run_test.py:70-72, n1/n2 = g1/g2 e^eta (1 + e^-eta
  (W - 1)) / W, to 5e-4 (the continuum terms move it by 1.25e-4, the J-bar
  quadrature by 3e-6; the reference's own bars are 3% mean, 4% max).
"""
import ctypes

import numpy as np
import pytest

from pubsed_amd import runtime, workloads as W
from oracle import oracle_py
from oracle.oracle_py import OracleBackend

H, KB, C, EV = 6.6260755e-27, 1.380658e-16, 2.99792458e10, 1.60217646e-12
R_P, T_S = 3.0e13, 5.0e4

THREE = dict(atom=dict(levels=((0.0, 2), (10.20, 8), (12.09, 18)), A=(4.696e8, 5.572e7, 4.408e7), chi=13.6),
             nu=(0.5e14, 2e16, 0), nu_log_del=0.0005, line_width=1e8)
TWO = dict(atom=dict(levels=((0.0, 2), (10.2, 8)), A=(4.7e8,), chi=60.0),
           nu=(2e14, 2e16, 0), nu_log_del=2e-4, line_width=5e7)


def three_level_table():
    import os
    return np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "three_level_nlte_analytic.dat"))


def two_level_radii():
    """zone centres of tests/2level_nlte/make_mod.py's 100 shells, 3e13..1e14 cm"""
    r_out = R_P + (1e14 - R_P) * np.arange(1, 101) / 100
    return r_out - 0.5 * (r_out[1] - r_out[0])


def dilution(r):
    return 0.5 * (1 - np.sqrt(1 - (R_P / r) ** 2))


def problem(space, spec, r, atom_file=None):
    """A pure-hydrogen zone per radius, at rest, T = 5e4 K, rho = 1e-20, with
    the spec's atom (or the atom of a Sedona atomic data file), grid and line
    width; returns (P, J_nu [nz][n_nu])."""
    nz = len(r)
    P = runtime.Problem(space, "c2_line", nz=nz, elements=((1, 1),), nu=spec["nu"],
                        nu_log_del=spec["nu_log_del"], fuzz=0, atom=spec["atom"], atom_file=atom_file)
    P.set("rho", np.full(nz, 1e-20))
    P.set("T", np.full(nz, T_S))
    P.op_params.omit_composition_decay = 1
    P.op_params.line_velocity_width = spec["line_width"]
    nuc = W.grid_center(P.nu_np)
    J = dilution(r)[:, None] * (2 * H * nuc ** 3 / C ** 2 / np.expm1(H * nuc / KB / T_S))[None, :]
    return P, np.ascontiguousarray(J)


def radiative_inputs(P, J, z):
    """The oracle's radiative rates of zone z: photoionisation I_i (Pic),
    recombination alpha_i (Rci, per n_e n_c), and per line (upper, lower)
    the ratio x = B_ul J-bar / A_ul and A_ul."""
    lib = oracle_py.load()
    A = P.atoms_np
    nl, nb = P.atoms.n_levels, len(A["bline_nu"])
    Pic, Rci, lJ = np.zeros(nl), np.zeros(nl), np.zeros(nb)
    Jz = np.ascontiguousarray(J[z])
    lib.or_nlte_radiative_rates(ctypes.byref(P.atoms), 0, ctypes.byref(P.nu), Jz.ctypes.data, T_S, 1,
                                float(P.op_params.line_velocity_width), Pic.ctypes.data, Rci.ctypes.data,
                                lJ.ctypes.data)
    x, Aul = {}, {}
    for l in range(nb):
        key = (int(A["bline_u"][l]), int(A["bline_l"][l]))
        x[key] = lJ[l] * A["bline_Bul"][l] / A["bline_A"][l]
        Aul[key] = A["bline_A"][l]
    return Pic, Rci, x, Aul


def closed_form_three(P, J, z):
    """problem_writeup.pdf eqs. (15)-(18): (n1/n2, n1/n3) from the rates,
    with s W / (e^zeta - 1) = B_ul J-bar / A_ul for each line."""
    Pic, Rci, x, Aul = radiative_inputs(P, J, z)
    g = P.atoms_np["lev_g"].astype(float)
    I1, I2, I3 = Pic[:3]
    a1, a2, a3 = Rci[:3]
    s = a1 + a2 + a3
    A21, A31, A32 = Aul[(1, 0)], Aul[(2, 0)], Aul[(2, 1)]
    x21, x31, x32 = x[(1, 0)], x[(2, 0)], x[(2, 1)]
    Aa = 1 + x31 + I3 / A31 * (1 - a3 / s) + A32 / A31 * (1 + x32)
    Bb = g[2] / g[0] * x31 + I1 / A31 * a3 / s
    Cc = g[2] / g[1] * A32 / A31 * x32 + I2 / A31 * a3 / s
    Dd = 1 + x21 + I2 / A21 * (1 - a2 / s) + g[2] / g[1] * A32 / A21 * x32
    Ee = g[1] / g[0] * x21 + I1 / A21 * a2 / s
    Ff = A32 / A21 * (1 + x32) + I3 / A21 * a2 / s
    num = Aa * Dd - Cc * Ff
    return num / (Aa * Ee + Bb * Ff), num / (Bb * Dd + Cc * Ee)


def closed_form_two(P, J, z):
    """Two levels plus continuum (writeup eq. (11) in terms of J-bar):
    n1/n2 = (1 + x + I2/A21 (1 - a2/(a1+a2))) / (g2/g1 x + I1/A21 a2/(a1+a2))."""
    Pic, Rci, x, Aul = radiative_inputs(P, J, z)
    g = P.atoms_np["lev_g"].astype(float)
    A21, x21 = Aul[(1, 0)], x[(1, 0)]
    f2 = Rci[1] / (Rci[0] + Rci[1])
    return (1 + x21 + Pic[1] / A21 * (1 - f2)) / (g[1] / g[0] * x21 + Pic[0] / A21 * f2)


def two_level_formula(r):
    """tests/2level_nlte/run_test.py:70-72"""
    eta = EV * 10.2 / KB / T_S
    Wd = dilution(r)
    return 2.0 / 8.0 * np.exp(eta) * (1 + np.exp(-eta) * (Wd - 1)) / Wd


def level_ratios(P, nz, cols):
    ln = P.get("lev_n").reshape(nz, -1)
    return [ln[:, 0] / ln[:, c] for c in cols]


@pytest.fixture(scope="module")
def three():
    tab = three_level_table()
    P, J = problem(runtime.Space(None), THREE, tab[:, 0])
    assert OracleBackend().opacity_build_nlte(P, J, np.ones(1, np.int32), use_collisions=0) == 0
    return P, J, tab


def test_three_level_closed_form(three):
    P, J, tab = three
    r12, r13 = level_ratios(P, len(tab), (1, 2))
    for z in range(len(tab)):
        c12, c13 = closed_form_three(P, J, z)
        assert abs(r12[z] / c12 - 1) <= 1e-12 and abs(r13[z] / c13 - 1) <= 1e-12, z


def test_three_level_analytic_table(three):
    P, J, tab = three
    r12, r13 = level_ratios(P, len(tab), (1, 2))
    outer = tab[:, 0] >= 2 * R_P
    for got, want in ((r12, tab[:, 1]), (r13, tab[:, 2])):
        e = np.abs(got / want - 1)
        assert e.max() <= 0.05 and e.mean() <= 0.01 and e[outer].max() <= 0.01
    # the ratios rise outward as the field dilutes
    assert np.all(np.diff(r12) > 0) and np.all(np.diff(r13) > 0)


def test_two_level_closed_forms():
    r = two_level_radii()
    P, J = problem(runtime.Space(None), TWO, r)
    assert OracleBackend().opacity_build_nlte(P, J, np.ones(1, np.int32), use_collisions=0) == 0
    (r12,) = level_ratios(P, len(r), (1,))
    for z in range(len(r)):
        assert abs(r12[z] / closed_form_two(P, J, z) - 1) <= 1e-12, z
    assert np.abs(r12 / two_level_formula(r) - 1).max() <= 5e-4


def test_three_level_atom_file_is_the_hand_built_atom(three):
    """tests/3level_nlte/param.lua:3 reads data/H_3lev_plusC_atomdata.hdf5
    (copied under tests/golden/atomdata/): read through pubsed_amd.atomdata it
    is the hand-built atom above -- the same levels, lines and derived line
    data, the bare stage's chi 999999 against 99999 (no stage lies above it,
    so nothing reads it) -- and the NLTE build from it returns the same
    populations and n_e."""
    import os
    P0, J0, tab = three
    path = os.path.join(os.path.dirname(__file__), "golden", "atomdata", "H_3lev_plusC_atomdata.hdf5")
    P, J = problem(runtime.Space(None), THREE, tab[:, 0], atom_file=path)
    a, b = P.atoms_np, P0.atoms_np
    for k in ("lev_off", "ion_off", "lev_E", "lev_g", "lev_ion", "bline_l", "bline_u", "bline_A", "bline_nu",
              "bline_Bul", "bline_Blu", "bline_f", "bline_bin"):
        assert np.array_equal(a[k], b[k]), k
    assert list(a["ion_chi"]) == [13.6, 999999.0] and list(b["ion_chi"]) == [13.6, 99999.0]
    assert OracleBackend().opacity_build_nlte(P, J, np.ones(1, np.int32), use_collisions=0) == 0
    for k in ("lev_n", "n_elec", "abs", "emis_cdf"):
        assert np.array_equal(P.get(k), P0.get(k)), k
