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
