"""Pins of the oracle's discrete diffusion Monte Carlo (transport_use_ddmc = 2;
transport::discrete_diffuse_DDMC and move_across_DDMC_interface,
src/transport/discrete_diffusion.cpp:111-509, hooked in as
transport.cpp:216-248, 272-309, 441-450).  CPU only.

The problem is the reference's diffusing pulse (tests/diffusing_pulse/models/
make_mod.py): 200 shells of 0.5 cm, rho = 1, grey kappa = 1 cm^2/g, pure
scattering, a uniform sphere of radius 5 cm of radiation at t = 0, static
(homologous with t_grid = 1e30 s), zone optical depth 0.5 > 0.1 threshold.

* Energy-weighted <r^2> = 3/5 R^2 + 6 D t with D = c / (3 kappa rho), and the
  enclosed-energy profile against the diffusion solution for a uniform sphere
  (Carslaw & Jaeger 10.3), both for DDMC and for plain Monte Carlo.
* Energy is conserved, and the J tally sums to E c t exactly.
* A thin envelope around the thick core: DDMC (leakage into Monte Carlo zones
  and conversion at the interface) and plain Monte Carlo give the same
  radial energy distribution within the Monte Carlo noise.
* Homologous flow: the adiabatic loss 1 - dt / t of the diffusing radiation.
* A threshold above every zone's optical depth leaves plain Monte Carlo
  untouched, draw for draw.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import erf

from pubsed_amd import runtime, abi
from oracle.oracle_py import OracleBackend

C = 2.99792458e10


def pulse(n, ddmc, t_end=3e-9, nz=200, dr=0.5, R=5.0, tau=0.1, nbin=1, seed=3, rho=None, t_grid=1e30,
          eps=0.0, device=None):
    P = runtime.Problem(runtime.Space(device), "c1_grey", nz=nz, grey=1.0, epsilon=eps, nu=(1e14, 1e16, nbin))
    r_out = dr * np.arange(1, nz + 1)
    P.set("r_out", r_out)
    P.set("vol", 4 * 3.14159 / 3 * (r_out ** 3 - np.concatenate([[0.0], r_out[:-1]]) ** 3))
    P.set("rho", np.ones(nz) if rho is None else rho)
    P.grid.r_out.min = 0.0
    P.grid.t_now = t_grid
    P.tparams.t_now = 0.0
    P.tparams.dt = t_end
    P.tparams.steady_state = 0
    P.tparams.use_ddmc = ddmc
    P.tparams.ddmc_tau_threshold = tau
    rng = np.random.default_rng(seed)
    r = R * rng.uniform(0, 1, n) ** (1 / 3)
    nh = rng.standard_normal((n, 3)); nh /= np.linalg.norm(nh, axis=1)[:, None]
    D = rng.standard_normal((n, 3)); D /= np.linalg.norm(D, axis=1)[:, None]
    x = r[:, None] * nh
    P.alloc_packets(n)
    P.load_packets(dict(x=x[:, 0], y=x[:, 1], z=x[:, 2], Dx=D[:, 0], Dy=D[:, 1], Dz=D[:, 2],
                        t=np.zeros(n), e=np.full(n, 1.0 / n), nu=np.full(n, 1e15),
                        ind=np.zeros(n, np.int32), type=np.zeros(n, np.int32),
                        fate=np.zeros(n, np.int32), draw=np.zeros(n, np.uint32),
                        id=np.arange(n, dtype=np.uint64)))
    return P


def run(P, n):
    ob = OracleBackend()
    ob.opacity_build(P)
    ob.wipe(P)
    ob.propagate(P, n)
    return P.packets_numpy(n)


def sphere_profile(r, t, a=5.0, D=C / 3):
    """Diffusion of a uniform sphere of radius a (unit density) in an infinite medium."""
    s = 2 * math.sqrt(D * t)
    return (0.5 * (erf((a - r) / s) + erf((a + r) / s))
            - math.sqrt(D * t / math.pi) / r * (math.exp(-(a - r) ** 2 / (4 * D * t))
                                                - math.exp(-(a + r) ** 2 / (4 * D * t))))


@pytest.mark.parametrize("ddmc", [0, 2])
def test_pulse_spreads_as_diffusion(ddmc):
    n, t = 30000, 3e-9
    pk = run(pulse(n, ddmc, t_end=t), n)
    assert np.all(pk["fate"] == abi.STOPPED)
    assert np.allclose(pk["t"], t, rtol=1e-12, atol=0)
    e = pk["e"]
    assert abs(e.sum() - 1) < 1e-12
    r2 = pk["x"] ** 2 + pk["y"] ** 2 + pk["z"] ** 2
    want = 0.6 * 25 + 6 * (C / 3) * t
    assert abs((e * r2).sum() / e.sum() / want - 1) < 0.03
    r = np.sqrt(r2)
    tot = quad(lambda q: sphere_profile(q, t) * 4 * math.pi * q * q, 0, 200)[0]
    for R in (3, 6, 10, 15, 20, 30):
        F = quad(lambda q: sphere_profile(q, t) * 4 * math.pi * q * q, 0, R)[0] / tot
        assert abs(e[r < R].sum() - F) < 0.012, (R, e[r < R].sum(), F)


def test_j_tally_is_energy_times_path():
    n, t = 5000, 3e-9
    P = pulse(n, 2, t_end=t)
    run(P, n)
    assert abs(P.get("J_nu").sum() / (C * t) - 1) < 1e-9
    assert P.get("e_abs").sum() == 0      # pure scattering


def test_ddmc_and_mc_agree_across_a_thick_thin_interface():
    """Thick core (rho = 1 to 12 cm, zone tau 0.5) in a thin envelope
    (rho = 1e-3, zone tau 5e-4 < 0.1): packets leak into Monte Carlo zones and
    are converted back at the interface."""
    n = 30000
    rho = np.where(0.5 * np.arange(1, 201) <= 12.0, 1.0, 1e-3)
    hs = []
    for ddmc in (0, 2):
        pk = run(pulse(n, ddmc, rho=rho), n)
        assert np.all(pk["fate"] == abi.STOPPED)
        r = np.sqrt(pk["x"] ** 2 + pk["y"] ** 2 + pk["z"] ** 2)
        h, _ = np.histogram(r, bins=[0, 4, 8, 12, 50, 100], weights=pk["e"])
        hs.append(h)
    sig = np.sqrt(hs[0] * (1 - hs[0]) / n) * math.sqrt(2)
    assert np.all(np.abs(hs[0] - hs[1]) < 4 * sig + 1e-3), (hs, sig)
    assert hs[1][3] + hs[1][4] > 0.5      # most of the energy left the core


@pytest.mark.parametrize("ddmc", [0, 2])
def test_adiabatic_loss_in_homologous_flow(ddmc):
    """t_grid = 1e-6 s: d ln rho / dt = 3 / t, so the comoving radiation energy
    falls by dt / t = 3e-3 over the step (v / c < 1e-2 everywhere it goes)."""
    n = 20000
    pk = run(pulse(n, ddmc, t_grid=1e-6), n)
    assert abs(pk["e"].sum() - (1 - 3e-9 / 1e-6)) < 3e-4


def test_threshold_above_every_zone_is_plain_monte_carlo():
    n = 4000
    a = run(pulse(n, 0, nbin=16), n)
    b = run(pulse(n, 2, nbin=16, tau=1.0), n)       # zone tau 0.5 <= 1
    for k in ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu", "ind", "fate"):
        assert np.array_equal(a[k], b[k]), k


def test_ddmc_with_frequency_bins_conserves_energy():
    """Non-grey grid: the effective inelastic scattering of the DDMC walk
    redistributes frequencies; with eps = 0 there is none, and the walk is the
    grey one."""
    n = 5000
    pk = run(pulse(n, 2, nbin=16), n)
    assert np.all(pk["fate"] == abi.STOPPED)
    assert abs(pk["e"].sum() - 1) < 1e-12
