"""Pins of the oracle's gas-state solve (Saha-Boltzmann LTE + Brent n_e),
radioactive decay and composition decay (CPU only).

Closed forms used:
* Boltzmann: n_i / n_j = g_i/g_j exp(-(E_i - E_j)/kT) inside one ion.
* 1-stage atom (neutral + the appended bare stage): ionisation balance
  n_e = n x1 with x1 = S/(n_e + S) gives n_e = (-S + sqrt(S^2 + 4 n S))/2,
  S = 2/lambda^3 (U_1/U_0) exp(-chi/kT); the Brent root must sit within the
  reference's tolerance (2 EPS |b| + tol/2, EPS = 3e-8) of it.
* 56Ni -> 56Co -> 56Fe: energy rate and composition against a numerical
  integration of the decay ODEs (scipy), not the Bateman formula itself.
"""
import ctypes
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from pubsed_amd import abi
from oracle import oracle_py

H, KB, KEV, ME, MP = 6.6260755e-27, 1.380658e-16, 8.6173324e-5, 9.10938188e-28, 1.67262158e-24
PI_REF = 3.14159


def make_atoms(atoms):
    """atoms: list of dict(levels=[(E, g, ion)], chi=[...]) (last ion = bare)."""
    lev_E, lev_g, lev_ion, chi, lo, io = [], [], [], [], [0], [0]
    for a in atoms:
        for E, g, ion in a["levels"]:
            lev_E.append(E); lev_g.append(g); lev_ion.append(ion)
        chi += a["chi"]
        lo.append(len(lev_E)); io.append(len(chi))
    arr = dict(lev_off=np.array(lo, np.int32), ion_off=np.array(io, np.int32),
               lev_E=np.array(lev_E), lev_g=np.array(lev_g, np.int32),
               lev_ion=np.array(lev_ion, np.int32), ion_chi=np.array(chi),
               zero=np.zeros(len(atoms), np.int32))
    A = abi.Atoms(len(atoms), int(np.max(np.diff(lo))), len(lev_E), len(chi),
                  *[arr[k].ctypes.data for k in ("lev_off", "ion_off", "lev_E", "lev_g", "lev_ion",
                                                 "ion_chi", "zero")],
                  0, None, None, None, None, None, None)
    return A, arr


def solve_state(A, dens, T, mf, elem_A):
    lib = oracle_py.load()
    n_lev, n_ion = A.n_levels, A.n_ions
    ip, fr, ln = np.zeros(n_ion), np.zeros(n_ion), np.zeros(n_lev)
    mf = np.asarray(mf, float)
    eA = np.asarray(elem_A, np.int32)
    mu = 1.0 / np.sum(mf / eA)
    ne = ctypes.c_double()
    nev = ctypes.c_int()
    st = lib.or_solve_state(ctypes.byref(A), len(mf), eA.ctypes.data, mf.ctypes.data, mu, dens, T,
                            ctypes.byref(ne), ip.ctypes.data, fr.ctypes.data, ln.ctypes.data,
                            ctypes.byref(nev))
    return st, ne.value, ip, fr, ln, nev.value


HYD = dict(levels=[(0.0, 2, 0), (10.2, 8, 0), (12.09, 18, 0), (0.0, 1, 1)], chi=[13.6, 99999.0])


@pytest.mark.parametrize("T,dens", [(1.0e4, 1e-13), (8.0e3, 1e-11), (2.0e4, 1e-14)])
def test_one_stage_saha_closed_form(T, dens):
    A, keep = make_atoms([HYD])
    st, ne, ip, fr, ln, nev = solve_state(A, dens, T, [1.0], [1])
    assert st == 0
    n = dens / (1 * MP)
    U0 = 2 + 8 * math.exp(-10.2 / KEV / T) + 18 * math.exp(-12.09 / KEV / T)
    lt = H * H / (2.0 * PI_REF * ME * KB * T)
    S = 2 / lt ** 1.5 * (1.0 / U0) * math.exp(-13.6 / KEV / T)
    ne_exact = (-S + math.sqrt(S * S + 4 * n * S)) / 2
    tol = 2 * 3e-8 * ne_exact + 0.5e-3
    assert abs(ne - ne_exact) <= 2 * tol, (ne, ne_exact)
    assert abs(fr.sum() - 1) < 1e-14
    assert abs(ip[0] - U0) <= 1e-14 * U0 and ip[1] == 1.0
    # Boltzmann ratios inside the neutral stage
    assert abs(ln[1] / ln[0] - 4 * math.exp(-10.2 / KEV / T)) <= 1e-13 * ln[1] / ln[0]
    # charge balance residual of the returned state
    assert abs(n * (ln[3] * 1) - ne) <= 1e-6 * ne + 1e-3


def test_multi_stage_invariants_and_levels():
    rng = np.random.default_rng(2)
    atoms = []
    for chis in ([7.9, 16.2, 30.7], [8.15, 16.35, 33.5]):
        levels = []
        for s, c in enumerate(chis):
            levels += [(0.0, int(rng.integers(1, 9)), s)] + [
                (float(E), int(rng.integers(1, 9)), s) for E in np.sort(rng.uniform(0, 0.8 * c, 6))]
        levels.append((0.0, 1, len(chis)))
        atoms.append(dict(levels=levels, chi=chis + [99999.0]))
    A, keep = make_atoms(atoms)
    st, ne, ip, fr, ln, nev = solve_state(A, 3e-13, 1.1e4, [0.4, 0.6], [56, 28])
    assert st == 0 and 5 <= nev <= 102
    for a in range(2):
        f = fr[keep["ion_off"][a]:keep["ion_off"][a + 1]]
        assert abs(f.sum() - 1) < 1e-13 and (f >= 0).all()
    # the returned n_e balances the charge of the final state within Brent's tolerance
    n = [3e-13 * 0.4 / (56 * MP), 3e-13 * 0.6 / (28 * MP)]
    q = 0.0
    for a in range(2):
        l0, l1 = keep["lev_off"][a], keep["lev_off"][a + 1]
        q += n[a] * np.sum(ln[l0:l1] * keep["lev_ion"][l0:l1])
    assert abs(q - ne) <= 4 * (2 * 3e-8 * ne + 0.5e-3) * max(1.0, abs(q - ne) / max(ne, 1)), (q, ne)


def test_brent_not_bracketed_status():
    """At 1000 K hydrogen is neutral even at n_e = 1e-10 n: f(a) < 0 too, so
    ne_brent_method flags error 1 (GasState.cpp:281-282)."""
    A, keep = make_atoms([HYD])
    st, ne, ip, fr, ln, nev = solve_state(A, 1e-10, 1000.0, [1.0], [1])
    assert st == 1


def decay_ode(t_end, X0_ni=1.0, X0_co=0.0):
    tni, tco = 757241.8, 9627379.
    f = lambda t, y: [-y[0] / tni, y[0] / tni - y[1] / tco, y[1] / tco]
    s = solve_ivp(f, (0, t_end), [X0_ni, X0_co, 0.0], rtol=1e-12, atol=1e-15, method="DOP853")
    return s.y[:, -1]


@pytest.mark.parametrize("t_day", [1.0, 20.0, 100.0])
def test_decay_against_ode(t_day):
    lib = oracle_py.load()
    t = t_day * 86400.0
    Z = np.array([28, 27, 26, 14], np.int32)
    A = np.array([56, 56, 56, 28], np.int32)
    X = np.array([0.7, 0.0, 0.0, 0.3])
    ni, co, fe = decay_ode(t)
    Xd = X.copy()
    lib.or_decay_composition(4, Z.ctypes.data, A.ctypes.data, Xd.ctypes.data, t)
    assert np.allclose(Xd[:3], 0.7 * np.array([ni, co, fe]), rtol=1e-9, atol=1e-14)
    assert abs(Xd.sum() - 1.0) < 1e-12
    # energy rate per gram: (N_ni/tau_ni E_ni + N_co/tau_co E_co), N per gram of initial Ni
    g = ctypes.c_double()
    L = lib.or_decay(4, Z.ctypes.data, A.ctypes.data, X.ctypes.data, t, ctypes.byref(g), 0)
    mev = 1.60217646e-6
    n0 = 0.7 / (56 * MP)
    want = n0 * (ni / 757241.8 * 1.725 * mev + co / 9627379. * 3.730 * mev)
    assert abs(L - want) <= 1e-9 * want
    gam = n0 * (ni / 757241.8 * 1.725 * mev + (1 - 0.0324) * co / 9627379. * 3.730 * mev)
    assert abs(g.value - gam / want) <= 1e-9


def test_rprocess_forced_heating_has_no_gamma_share():
    lib = oracle_py.load()
    Z = np.array([28], np.int32)
    A = np.array([56], np.int32)
    X = np.array([1.0])
    g = ctypes.c_double(-1)
    L = lib.or_decay(1, Z.ctypes.data, A.ctypes.data, X.ctypes.data, 5 * 86400.0, ctypes.byref(g), 1)
    td = 5.0
    eps = 8.4939e9 * td ** -1.3642 + 8.3425e9 * math.exp(-td / 3.628) + 8.8616e8 * math.exp(-td / 10.847)
    x = 2 * 0.17 * td ** 0.74
    f = 0.36 * (math.exp(-0.56 * td) + math.log(1 + x) / x)
    assert abs(L - eps * f) <= 1e-12 * L and g.value == 0.0
