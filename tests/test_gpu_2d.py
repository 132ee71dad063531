"""The 2D axisymmetric (varpi, z) configuration on the GPU against the CPU
oracle (grid_2D_cyln.cpp:315-569, hydro_homologous.cpp:27-61,
transport.cpp:26-484): the homologous step, the opacity build, emission,
FP64 propagation path for path, FP32-geometry statistics, and a
time-dependent run of several steps whose census packets persist.

Sizes are reduced so the oracle finishes in seconds (16 x 32 zones, 512
frequency bins, 4000 fuzz lines); the full 128 x 256 x 4096 shape runs in
bench.py --config c4_2d.
"""
import numpy as np
import pytest

from pubsed_amd import runtime, abi, workloads as W
from oracle.oracle_py import OracleBackend

pytestmark = pytest.mark.gpu

SMALL = dict(nx=16, nzz=32, nu=(W.C_LIGHT / 1e-4, W.C_LIGHT / 1e-5, 512), n_lines=4000)


def pair(gpu, n, **over):
    cfg = dict(SMALL)
    cfg.update(over)
    Pg = runtime.Problem(runtime.Space("cuda:0"), "c4_2d", **cfg)
    Ph = runtime.Problem(runtime.Space(None), "c4_2d", **cfg)
    Pg.alloc_packets(n)
    Ph.alloc_packets(n)
    return Pg, Ph, OracleBackend()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def frac_diff(g, h):
    g, h = np.asarray(g), np.asarray(h)
    tot = np.abs(g).sum() + np.abs(h).sum()
    return float(np.abs(g - h).sum() / tot) if tot > 0 else 0.0


def test_homologous_step_2d(gpu):
    Pg, Ph, ob = pair(gpu, 16)
    dt = 0.3 * W.DAY
    print("case homologous 2d", flush=True)
    for _ in range(2):
        gpu.homologous_step(Pg, dt)
        ob.homologous_step(Ph, dt)
    gpu.sync()
    for k in ("x_out", "z_out", "rho", "vol"):
        assert rel(Pg.get(k), Ph.get(k)) <= 1e-15, k
    for gk in ("x_out", "z_out"):
        lg, lh = getattr(Pg.grid, gk), getattr(Ph.grid, gk)
        assert (lg.min, lg.del_, lg.n) == (lh.min, lh.del_, lh.n)
    assert Pg.grid.t_now == Ph.grid.t_now == (Ph.ej["t"] + dt) + dt
    # mass is conserved by the expansion
    m0 = np.sum(Ph.ej["rho"] * Ph.ej["vol"])
    assert abs(np.sum(Ph.get("rho") * Ph.get("vol")) / m0 - 1) <= 1e-12


def test_opacity_build_2d(gpu):
    Pg, Ph, ob = pair(gpu, 16)
    print("case opacity 2d", flush=True)
    gpu.opacity_build(Pg)
    gpu.sync()
    ob.opacity_build(Ph)
    for k in ("abs", "scat", "emis_cdf", "eps_imc", "L_thermal", "n_elec"):
        g, h = Pg.get(k), Ph.get(k)
        assert np.max(np.abs(g - h)) <= 1e-10 * np.max(np.abs(h)) + 1e-300, k


def test_emission_2d(gpu):
    n = 4096
    Pg, Ph, ob = pair(gpu, n)
    gpu.opacity_build(Pg)
    ob.opacity_build(Ph)
    gpu.wipe(Pg)
    ob.wipe(Ph)
    print("case emit 2d", flush=True)
    assert gpu.emit(Pg, n, id_base=3) == n
    gpu.sync()
    assert ob.emit(Ph, n, id_base=3) == n
    a, b = Pg.packets_numpy(n), Ph.packets_numpy(n)
    for k in ("ind", "type", "draw", "id"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu"):
        assert rel(a[k], b[k]) <= 1e-12, k
    # every packet lies in the zone it was emitted in
    nzz = Ph.ej["nzz"]
    p = np.sqrt(b["x"] ** 2 + b["y"] ** 2)
    ix = np.searchsorted(Ph.get("x_out"), p, side="right")
    iz = np.searchsorted(Ph.get("z_out"), b["z"], side="right")
    assert np.mean(ix * nzz + iz == b["ind"]) >= 0.999


@pytest.mark.parametrize("steady", [1, 0])
def test_propagate_fp64_2d_matches_oracle(gpu, steady):
    n = 3000
    Pg, Ph, ob = pair(gpu, n, steady=steady, dt=0.3 * W.DAY)
    for P, be in ((Pg, gpu), (Ph, ob)):
        be.opacity_build(P)
        be.wipe(P)
        be.emit(P, n)
    cnt = Pg.space.zeros(10, "i64")
    print("case propagate fp64 2d steady=%d" % steady, flush=True)
    gpu.propagate(Pg, n, cnt)
    gpu.sync()
    ob.propagate(Ph, n)
    a, b = Pg.packets_numpy(n), Ph.packets_numpy(n)
    same = (a["fate"] == b["fate"]) & (a["ind"] == b["ind"]) & (a["draw"] == b["draw"])
    print("identical paths", same.mean(), "fates", np.bincount(b["fate"], minlength=4))
    assert same.mean() >= 0.999
    if not steady:
        assert np.sum(b["fate"] == abi.STOPPED) > 0
        st = b["fate"] == abi.STOPPED
        assert np.allclose(a["t"][st & same], b["t"][st & same], rtol=1e-12, atol=0)
    # the escape spectrum in the 2D benchmark's 10 mu bins (and 0.1-day time
    # bins when time-dependent), bin for bin
    assert Pg.spec_shape[2] == 10 and (steady or Pg.spec_shape[0] == 100)
    for k in ("e_abs", "J_nu", "L_radio_dep", "opt_flux", "opt_click"):
        d = frac_diff(Pg.get(k), Ph.get(k))
        assert d <= 10 * (1 - same.mean()) + 1e-9, (k, d)


def test_propagate_mixed_2d_statistics(gpu):
    n = 20000
    Pg, Ph, ob = pair(gpu, n, steady=0, dt=0.3 * W.DAY)
    for P, be in ((Pg, gpu), (Ph, ob)):
        be.opacity_build(P)
        be.wipe(P)
        be.emit(P, n)
    Pg.tparams.precision = abi.PREC_MIXED
    print("case propagate mixed 2d", flush=True)
    gpu.propagate(Pg, n)
    gpu.sync()
    ob.propagate(Ph, n)
    fg, fh = Pg.packets_numpy(n)["fate"], Ph.packets_numpy(n)["fate"]
    for f in (abi.ESCAPED, abi.STOPPED):
        pg, ph = (fg == f).mean(), (fh == f).mean()
        sig = np.sqrt(pg * (1 - pg) / n + ph * (1 - ph) / n)
        assert abs(pg - ph) <= 5 * sig + 1e-12, (f, pg, ph, sig)
    for k in ("e_abs", "L_radio_dep"):
        g, h = Pg.get(k).sum(), Ph.get(k).sum()
        assert abs(g - h) <= 0.05 * abs(h), (k, g, h)


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_time_dependent_census_2d(gpu, precision):
    """Three steps of homologous expansion + transport; packets still moving
    at the end of a step persist into the next.  FP64: census sets and
    tallies follow the oracle; FP32 geometry: census counts and deposition
    within packet noise."""
    # FP64 follows the oracle path by path; FP32 is compared through totals,
    # whose packet noise (measured on the oracle with different Philox keys:
    # 0.54 % for e_abs at 8000 packets per step) needs the larger count
    n_emit, steps = (2000 if precision == abi.PREC_FP64 else 8000), 3
    cap = n_emit * steps
    dt = 0.5 * W.DAY
    Pg, Ph, ob = pair(gpu, cap, dt=dt)
    Pg.tparams.precision = precision
    scr = Pg.alloc_packets(cap, name="scr")
    for s in range(steps):
        print("case census 2d precision=%d step %d" % (precision, s), flush=True)
        got_g, alive_g, esc_g = runtime.time_step(gpu, Pg, dt, n_emit, scr, s)
        gpu.sync()
        got_h, alive_h, esc_h = runtime.time_step(ob, Ph, dt, n_emit, None, s)
        assert got_g == got_h == n_emit
        assert Pg.grid.t_now == Ph.grid.t_now and Pg.t_now == Ph.t_now
        a, b = Pg.packets_numpy(alive_g), Ph.packets_numpy(alive_h)
        # every survivor stopped at the end of the step
        assert np.all(a["fate"] == abi.STOPPED) and np.all(b["fate"] == abi.STOPPED)
        assert np.allclose(b["t"], Ph.t_now, rtol=1e-12, atol=0)
        print("alive", alive_g, alive_h, "escaped", esc_g, esc_h)
        if precision == abi.PREC_FP64:
            common = np.intersect1d(a["id"], b["id"]).size
            assert common >= 0.99 * max(alive_g, alive_h)
            for k in ("e_abs", "J_nu", "L_radio_dep"):
                assert frac_diff(Pg.get(k), Ph.get(k)) <= 0.02, k
        else:
            tot = alive_g + esc_g
            pg, ph = alive_g / tot, alive_h / (alive_h + esc_h)
            sig = np.sqrt(pg * (1 - pg) / tot + ph * (1 - ph) / tot)
            assert abs(pg - ph) <= 5 * sig + 1e-12
            g, h = Pg.get("e_abs").sum(), Ph.get("e_abs").sum()
            assert abs(g - h) <= 5 * 1.414 * 0.0054 * abs(h), (g, h)
