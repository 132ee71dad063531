"""The host C++ drop-in (pubsed_amd/host/sedona_transport.cpp, a transport
class with the reference's step order) drives the same kernels as the Python
binding: one grey 1D step through it must reproduce the Python path's
deposition, escape count and escaping flux for the same seed."""
import json
import os
import subprocess

import numpy as np
import pytest

from pubsed_amd import runtime, abi
from pubsed_amd.build import DRIVER

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_host_driver_matches_python_path(gpu, precision):
    n = 20000
    assert os.path.exists(DRIVER), "build the host driver first"
    print("case host driver", precision, flush=True)
    r = subprocess.run([DRIVER, str(n), str(precision)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    P = runtime.Problem(runtime.Space("cuda:0"), "c1_grey")
    P.tparams.precision = precision
    P.alloc_packets(n)
    gpu.opacity_build(P)
    gpu.wipe(P)
    gpu.emit(P, n)
    gpu.propagate(P, n)
    fates = P.packets_numpy(n)["fate"]
    L_dep = P.get("L_radio_dep").sum()          # before normalisation: erg/s (dt = 1)
    flux = P.get("opt_flux").sum()
    gpu.sync()
    assert got["escaped"] == int((fates == abi.ESCAPED).sum())
    assert got["alive"] == 0
    assert abs(got["L_dep"] / L_dep - 1) < 1e-9
    assert abs(got["optical_flux"] / flux - 1) < 1e-9


def test_host_driver_writes_reference_text_files(gpu, tmp_path):
    """transport::output_spectrum and write_plotfile through the host class:
    spectrum_0.dat / gamma_0.dat hold the Python path's escaping flux per unit
    frequency (1024 rows: the single time, mu and phi bins add no column), and
    plt_00000.dat the zone state."""
    from pubsed_amd import output as O
    n = 20000
    print("case host driver output", flush=True)
    r = subprocess.run([DRIVER, str(n), str(abi.PREC_FP64), str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    P = runtime.Problem(runtime.Space("cuda:0"), "c1_grey")
    P.tparams.precision = abi.PREC_FP64
    P.alloc_packets(n)
    gpu.opacity_build(P)
    gpu.wipe(P)
    gpu.emit(P, n)
    gpu.propagate(P, n)
    gpu.sync()
    g = P.spec_grids["nu"]
    x = np.asarray(g["x"]); dnu = x - np.concatenate([[g["min"]], x[:-1]])
    for stem, key in (("spectrum_0", "opt"), ("gamma_0", "gam")):
        counts, rows = O.read_spectrum_dat(str(tmp_path / (stem + ".dat")))
        assert counts == (1, 1024, 1, 1) and rows.shape == (1024, 3)
        want = P.get(key + "_flux") / dnu
        assert np.array_equal(rows[:, 2], P.get(key + "_click"))
        hit = want > 0
        # escaping gamma-rays carry nu in MeV, off the optical grid: bin 0
        assert hit.sum() > (10 if key == "opt" else 0), stem
        assert np.all(rows[~hit, 1] == 0)
        assert np.allclose(rows[hit, 1], want[hit], rtol=2e-5, atol=0), stem
    plt = np.loadtxt(str(tmp_path / "plt_00000.dat"))
    assert plt.shape == (100, 8 + 4)
    assert np.allclose(plt[:, 3], P.get("T"), rtol=1e-8)
    assert np.allclose(plt[:, 1], P.get("rho"), rtol=1e-8)


@pytest.mark.parametrize("precision", [abi.PREC_FP64, abi.PREC_MIXED])
def test_host_checkpoint_restart(gpu, tmp_path, precision):
    """transport::write_checkpoint / read_checkpoint: three time-dependent
    steps straight through and 1 + checkpoint + 2 in a fresh transport end
    with the same census, escapes, clock and tallies (up to the order of the
    atomic additions); the checkpoint the C++ class wrote loads into the
    Python problem of the same grid through pubsed_amd.checkpoint."""
    from pubsed_amd import checkpoint
    n = 20000
    chk = str(tmp_path / "chk")
    print("case host checkpoint", precision, flush=True)
    r = subprocess.run([DRIVER, str(n), str(precision), "", chk], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    a, c = got["straight"], got["restarted"]
    print(a, c)
    assert a["alive"] == c["alive"] > 0 and a["escaped"] == c["escaped"] and a["t_now"] == c["t_now"]
    for k in ("L_dep", "E_abs_rate", "J_sum", "spec_flux", "spec_flux_weighted"):
        assert abs(c[k] / a[k] - 1) < 1e-10, k
    # the spectrum (half-day time bins from 18 d, 4 mu bins) holds the
    # escapes of all three steps in both runs: the restarted run's first-step
    # escapes come from the checkpoint
    assert a["spec_clicks"] == c["spec_clicks"] > 0
    assert a["spec_bins"] == c["spec_bins"] == 20 * 1024 * 4
    P = runtime.Problem(runtime.Space(None), "c1_grey", steady=0, spec_t0=18 * 86400.0, spec_t1=28 * 86400.0,
                        spec_dt=0.5 * 86400.0, spec_n_mu=4)
    P.alloc_packets(3 * n)
    meta = checkpoint.read_checkpoint(P, chk)
    assert P.n_packets == meta["n_packets"] > 0
    pk = P.packets_numpy(P.n_packets)
    assert np.unique(pk["id"]).size == P.n_packets
    # census packets stop at the step's end (the mixed form's FP32 flight
    # distance puts them there to FP32 accuracy)
    tol = 1e-12 if precision == abi.PREC_FP64 else 1e-6
    assert np.all(pk["fate"] == abi.STOPPED) and np.allclose(pk["t"], meta["t_now"], rtol=tol, atol=0)
    assert np.all((pk["ind"] >= 0) & (pk["ind"] < P.nz))
    spec = np.load(os.path.join(chk, "state", "opt_click.npy"))
    assert spec.shape == (20 * 1024 * 4,) and spec.sum() > 0
