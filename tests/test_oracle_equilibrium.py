"""Pins of the oracle's radiative-equilibrium temperature
(transport::solve_eq_temperature, src/transport/solve_equilibrium.cpp:66-320;
CPU only).

* Stefan-Boltzmann: a grey absorption coefficient on a wide frequency grid
  emits sum 4 pi kappa B_nu(T) dnu -> 4 kappa sigma T^4, so e_abs =
  4 kappa sigma T0^4 returns T0 (to the grid's quadrature error);
* the one-bin (grey) branch: 4 pi kappa (sigma / pi) T^4 inverted exactly;
* a zone bathed in a blackbody field: e_abs = sum 4 pi kappa_nu B_nu(T0) dnu
  for an arbitrary kappa_nu returns T0 to the Brent tolerance (1e-2 K);
* transport_set_Tgas_to_Trad: T = (e_rad / a)^(1/4).
"""
import ctypes
import math

import numpy as np
import pytest

from pubsed_amd import abi, runtime, workloads as W
from oracle import oracle_py

H, KB, C, SB, A_RAD = 6.6260755e-27, 1.380658e-16, 2.99792458e10, 5.6704e-5, 7.5657e-15


def solve(nu, abs_tab, e_abs, e_rad=None, tmin=1000.0, tmax=1e8, trad=0):
    lib = oracle_py.load()
    L, keep = runtime.locate_struct(runtime.Space(None), nu)
    nz = len(e_abs)
    T = np.zeros(nz)
    ne = np.zeros(nz, np.int32)
    e_rad = np.zeros(nz) if e_rad is None else np.asarray(e_rad, float)
    lib.or_solve_eq_temperature(ctypes.byref(L), nz, np.ascontiguousarray(abs_tab).ctypes.data,
                                np.ascontiguousarray(e_abs, float).ctypes.data, e_rad.ctypes.data, tmin, tmax,
                                trad, T.ctypes.data, ne.ctypes.data)
    return T, ne


@pytest.mark.parametrize("T0", [3.0e3, 1.2e4, 5.0e4])
def test_grey_absorber_obeys_stefan_boltzmann(T0):
    nu = W.nu_log_grid(1e11, 1e18, 3000)
    kappa = 2.5e-12
    e_abs = [4 * kappa * SB * T0 ** 4]
    T, ne = solve(nu, np.full(nu["n"], kappa), e_abs)
    assert abs(T[0] / T0 - 1) < 2e-4, T
    assert ne[0] < 100


def test_one_bin_branch_inverts_exactly():
    nu = dict(n=1, type=abi.LOC_NONE, min=1e14, del_=0.0, x=np.array([1e16]))
    kappa = 3e-11
    T0 = np.array([2000.0, 7777.0, 31000.0])
    e_abs = 4 * kappa * SB * T0 ** 4
    T, _ = solve(nu, np.full((3, 1), kappa), e_abs)
    # Brent stops within tol1 = 2 EPS T + 0.5e-2 of the root
    assert np.all(np.abs(T - T0) <= 2 * 3e-8 * T0 + 0.5e-2 + 1e-6), T - T0


def test_blackbody_field_returns_its_temperature():
    nu = W.nu_log_grid(1e13, 1e17, 1024)
    rng = np.random.default_rng(3)
    nz = 6
    kap = 10.0 ** rng.uniform(-14, -9, (nz, nu["n"]))
    T0 = np.array([2.0e3, 4.0e3, 8.0e3, 1.5e4, 3.0e4, 6.0e4])
    nuc, dnu = W.grid_center(nu), W.grid_delta(nu)
    with np.errstate(over="ignore"):
        e_abs = np.array([np.sum(4 * math.pi * kap[i] * 2 * H * nuc ** 3 / C ** 2 / np.expm1(H * nuc / KB / T0[i]) * dnu)
                          for i in range(nz)])
    T, ne = solve(nu, kap, e_abs)
    # the reference's pi is 3.14159 in 4 pi; numpy's 4 pi differs by 8.4e-7
    # relative, which moves T by 2e-7 T
    assert np.all(np.abs(T - T0) <= 1e-2 + 5e-7 * T0), T - T0


def test_set_tgas_to_trad():
    nu = W.nu_log_grid(1e13, 1e17, 16)
    e_rad = np.array([1.0, 7.5, 3e3])
    T, _ = solve(nu, np.ones((3, 16)), np.zeros(3), e_rad=e_rad, trad=1)
    assert np.allclose(T, (e_rad / A_RAD) ** 0.25, rtol=1e-15, atol=0)
