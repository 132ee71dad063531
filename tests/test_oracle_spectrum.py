"""The escape spectrum (spectrum_array::count, spectrum_array.cpp:215-246,
laid out [t][nu][mu][phi] by spectrum_array::index) and its life over the
steps of a run, on the oracle (CPU):

* mu / phi / time binning: the click counts of a 2D time-dependent step with
  10 mu bins (the reference's 2D benchmark,
  tests/ellipsoidal_outflow_2D/2D_benchmark/param.lua:48), 3 phi bins and
  0.1-day time bins equal a numpy histogram of the escaped packets' final
  (t_obs, nu, D_z, atan2(D_y, D_x)) on the same edges, packet for packet
  (a transposed index or a mu taken from another component fails it);
  the flux sums over mu and phi to the single-bin spectrum of the same run.
* A time-dependent run keeps every step's escapes in its spectrum (the
  reference wipes spectra only in steady iteration, SedonaClass.cpp:349):
  three steps with the spectrum read and cleared after each add up to the
  spectrum of three steps run without clearing.
"""
import math

import numpy as np

from pubsed_amd import abi, runtime, workloads as W
from oracle.oracle_py import OracleBackend

# the 2D grid with a grey absorber (kappa 0.05 cm^2/g, epsilon 0.1) so that a
# 5-day step lets photons escape
SMALL2D = dict(nx=8, nzz=16, steady=0, radiative_eq=0, grey=0.05, epsilon=0.1, escat=0, fuzz=0)


def edges(g):
    return np.concatenate([[g["min"]], np.asarray(g["x"])])


def test_mu_phi_time_bins_match_a_histogram_of_the_escapes():
    n, dt = 6000, 5 * W.DAY
    ob = OracleBackend()
    runs = {}
    for n_mu, n_phi in ((10, 3), (1, 1)):
        P = runtime.Problem(runtime.Space(None), "c4_2d", spec_n_mu=n_mu, spec_n_phi=n_phi, **SMALL2D)
        P.alloc_packets(n)
        P.tparams.dt = P.op_params.dt = dt
        ob.opacity_build(P)
        ob.wipe(P, 1)
        ob.emit(P, n)
        ob.propagate(P, n)
        runs[n_mu] = P
    P = runs[10]
    shape = P.spec_shape
    assert shape[2:] == (10, 3) and shape[0] > 1
    pk = P.packets_numpy(n)
    esc = (pk["fate"] == abi.ESCAPED) & (pk["type"] == abi.PHOTON)
    assert esc.sum() > 500
    t, nu, mu = pk["t"][esc], pk["nu"][esc], pk["Dz"][esc]
    phi = np.arctan2(pk["Dy"][esc], pk["Dx"][esc])
    phi = np.where(phi < 0, phi + 2 * W.PI_REF, phi)
    g = P.spec_grids
    nn = shape[1]
    # off-grid frequencies are counted in bin 0; off-grid t / mu / phi dropped
    ib = np.searchsorted(np.asarray(g["nu"]["x"]), nu, side="right")
    ib = np.where((nu < g["nu"]["min"]) | (ib >= nn), 0, ib)
    h, _ = np.histogramdd(np.stack([t, ib + 0.5, mu, phi], axis=1),
                          bins=[edges(g["time"]), np.arange(nn + 1), edges(g["mu"]), edges(g["phi"])])
    clicks = P.get("opt_click").reshape(shape)
    assert clicks.sum() > 0.9 * esc.sum()
    assert np.array_equal(clicks, h)
    # at least 8 of the 10 mu bins and every phi bin are populated
    assert (clicks.sum(axis=(0, 1, 3)) > 0).sum() >= 8 and np.all(clicks.sum(axis=(0, 1, 2)) > 0)
    # the same packets in one mu / phi bin: the flux summed over mu and phi
    one = runs[1].get("opt_flux").reshape(runs[1].spec_shape)[..., 0, 0]
    assert np.allclose(P.get("opt_flux").reshape(shape).sum(axis=(2, 3)), one, rtol=1e-12, atol=0)


def test_time_dependent_spectrum_accumulates_over_steps():
    n_emit, dt, steps = 1500, 2 * W.DAY, 3
    ob = OracleBackend()
    A = runtime.Problem(runtime.Space(None), "c1_grey", nz=24, steady=0, radiative_eq=0)
    B = runtime.Problem(runtime.Space(None), "c1_grey", nz=24, steady=0, radiative_eq=0)
    per_step = []
    for P in (A, B):
        P.alloc_packets(steps * n_emit)
    for s in range(steps):
        runtime.time_step(ob, A, dt, n_emit, None, s)
        runtime.time_step(ob, B, dt, n_emit, None, s)
        per_step.append((B.get("opt_flux"), B.get("opt_click")))
        for k in ("opt_flux", "opt_click", "gam_flux", "gam_click"):
            B.set(k, np.zeros(B.n_spec))
    assert all(c.sum() > 0 for _, c in per_step)
    assert np.array_equal(A.get("opt_click"), sum(c for _, c in per_step))
    assert np.allclose(A.get("opt_flux"), sum(f for f, _ in per_step), rtol=1e-12, atol=0)
