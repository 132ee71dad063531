"""Pins of the oracle's bound-bound (Voigt-profile line) opacity
(GasState::bound_bound_opacity, GasState_opacities.cpp:83-91,379-408, over
AtomicSpecies::bound_bound_opacity, AtomicSpecies_opacities.cpp:224-282),
one zone, the hydrogen-like atom, one line between its Boltzmann levels:

* Kirchhoff: the line's emissivity over its opacity is, bin by bin,
  2 h nu^3 / c^2 r / (1 - r) with r = n_u g_l / (n_l g_u) (nu the bin
  centre): the extinction carries (1 - r) and the emission n_u A h / 4 pi.
  In LTE r = e^{-dE / kT} and this is B_nu at the line's excitation
  temperature (the populations take the exponent in eV with k in eV/K,
  6.7e-6 apart from h nu_0 / kT with k in erg/K).
* Strength: on a grid that resolves the profile, sum_j kappa_j nu_j^2 dnu_j /
  nu_0^2 is the classical integrated line opacity sigma_tot f_lu n_l
  (1 - r) with f_lu from A_ul by the Einstein relation
  (lambda^2 A g_u / g_l / (8 pi sigma_tot)); the profile's +-5 Doppler
  widths hold all but erfc(5) of a Gaussian core.
* A line narrower than its bin (+-5 Doppler widths inside one bin) adds
  nothing: the reference sums over [locate(nu_0 - 5 dnu), locate(nu_0 + 5
  dnu)), which is empty then.
"""
import math

import numpy as np
import pytest

from test_oracle_opacity import OneZone, bin_emissivity, centers, H, KB, C, MP
from test_oracle_state import HYD

EV_ERG = 1.60217646e-12
SIGMA_TOT = 0.0265400193567
PI_REF = 3.14159


def add_bb_line(z, l, u, A_ul, width):
    """One level-resolved line l -> u with Einstein A of the OneZone's atom;
    bound-bound opacity on, every other term off; line_velocity_width."""
    E = z.akeep["lev_E"]
    nn = z.nu["n"]
    nu = np.array([(E[u] - E[l]) * EV_ERG / H])
    b = np.minimum(np.searchsorted(z.nu["x"], nu, side="right"), nn - 1).astype(np.int32)
    off = np.zeros(nn + 1, np.int64)
    off[1:] = np.cumsum(np.bincount(b, minlength=nn))
    ll, lu = np.array([l], np.int32), np.array([u], np.int32)
    A = np.array([A_ul])
    zero = np.zeros(1)
    z.bb = dict(ll=ll, lu=lu, nu=nu, A=A, b=b, off=off, zero=zero)
    at = z.A
    at.n_blines = 1
    at.bline_l, at.bline_u, at.bline_nu = ll.ctypes.data, lu.ctypes.data, nu.ctypes.data
    at.bline_f = at.bline_Bul = at.bline_Blu = zero.ctypes.data
    at.bline_A = A.ctypes.data
    at.bline_bin, at.bline_bin_off = b.ctypes.data, off.ctypes.data
    z.p.use_bound_bound = 1
    z.p.line_velocity_width = width
    return nu[0]


@pytest.mark.parametrize("T", [6.0e3, 1.2e4, 3.0e4])
def test_bound_bound_obeys_kirchhoff(T):
    z = OneZone([], T=T, dens=1e-13)
    nu0 = add_bb_line(z, 0, 1, 4.7e8, 3e8)
    o = z.run()
    kap = o["abs"]
    on = kap > 0
    assert on.sum() >= 5
    emis, _ = bin_emissivity(z, o)
    nuc = centers(z)
    n = o["lev_n"]
    r = n[1] * 2.0 / (n[0] * 8.0)
    assert abs(r / math.exp(-10.2 / 8.6173324e-5 / T) - 1) < 1e-12
    want = 2 * H * nuc ** 3 / C ** 2 * r / (1 - r)
    assert np.allclose(emis[on] / kap[on], want[on], rtol=1e-9)


def test_bound_bound_integrated_strength():
    T, dens = 1.0e4, 1e-13
    z = OneZone([], T=T, dens=dens, grid=(2.2e15, 2.8e15, 6000))
    A_ul = 4.7e8
    nu0 = add_bb_line(z, 0, 1, A_ul, 3e8)
    o = z.run()
    x = z.nu["x"]
    dnu = np.diff(np.concatenate([[z.nu["min"]], x]))
    nuc = centers(z)
    got = np.sum(o["abs"] * nuc ** 2 * dnu) / nu0 ** 2
    gl, gu = 2.0, 8.0
    lam = C / nu0
    f_lu = lam * lam * A_ul * gu / gl / (8 * PI_REF * SIGMA_TOT)
    n_l = o["lev_n"][0] * dens / MP
    r = o["lev_n"][1] * gl / (o["lev_n"][0] * gu)
    want = SIGMA_TOT * f_lu * n_l * (1 - r)
    assert abs(got / want - 1) < 1e-3, (got, want)


def test_line_narrower_than_its_bin_adds_nothing():
    z = OneZone([], T=1.0e4, dens=1e-13)
    add_bb_line(z, 0, 1, 4.7e8, 0.0)          # thermal width: 4e-5 of nu_0, bins of 1.2 %
    o = z.run()
    assert np.all(o["abs"] == 0)
    z2 = OneZone([], T=1.0e4, dens=1e-13)
    add_bb_line(z2, 0, 1, 4.7e8, 3e8)
    assert np.count_nonzero(z2.run()["abs"]) >= 5


@pytest.mark.parametrize("T", [8.0e3, 2.0e4])
def test_minimum_extinction_cut(T):
    """opacity_minimum_extinction (AtomicSpecies_opacities.cpp:259): a line
    is skipped when alpha_0 / (nu_0^2 dnu) falls below the cut, with
    alpha_0 = n_l n g_u/g_l A c^2 / (8 pi) (1 - r) and dnu = (v / c) nu_0.
    The threshold is computed here from the state the build returns; a cut
    0.1 % above it removes the line, 0.1 % below leaves the tables as with
    no cut (a threshold formula off by a factor nu_0, dnu or the stimulated
    term fails one side)."""
    width, A_ul, dens = 3e8, 4.7e8, 1e-13
    z0 = OneZone([], T=T, dens=dens)
    nu0 = add_bb_line(z0, 0, 1, A_ul, width)
    o0 = {k: v.copy() for k, v in z0.run().items()}
    assert np.count_nonzero(o0["abs"]) >= 5
    gl, gu = 2.0, 8.0
    nl, nup = o0["lev_n"][0], o0["lev_n"][1]
    n_dens = dens / MP
    alpha_0 = nl * n_dens * gu / gl * A_ul / (8 * PI_REF) * C * C * (1 - nup * gl / (nl * gu))
    thr = alpha_0 / (nu0 * nu0 * (width / C * nu0))
    for fac, kept in ((0.999, True), (1.001, False)):
        z = OneZone([], T=T, dens=dens)
        add_bb_line(z, 0, 1, A_ul, width)
        z.p.minimum_extinction = fac * thr
        o = z.run()
        if kept:
            assert np.array_equal(o["abs"], o0["abs"]) and np.array_equal(o["emis_cdf"], o0["emis_cdf"])
        else:
            assert np.all(o["abs"] == 0)
