"""GPU parity of the inner boundary source (sgpu_emit_inner_source;
transport::emit_inner_source, emission.cpp:370-480) against the oracle,
packet for packet, and the reference's spherical lightbulb
(tests/spherical_lightbulb/1D) on the GPU at 5e7 packets with its own
criteria over every bin: T_rad and T_gas within 1 % of the dilution
temperature outside the core, the escaping spectrum and zone 50's J_nu
within a mean 10 % of the (diluted) Planck function."""
import numpy as np
import pytest

from pubsed_amd import abi, runtime
from oracle.oracle_py import OracleBackend
from test_oracle_lightbulb import lightbulb, lightbulb_run, check_lightbulb, core

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("r_core,freq", [(5.0e14, 0.0), (0.0, 0.0), (5.0e14, 7.7e14)])
def test_inner_source_matches_oracle(gpu, r_core, freq):
    n = 50000
    Pg, _ = lightbulb(runtime.Space("cuda:0"), n + 7)
    Ph, _ = lightbulb(runtime.Space(None), n + 7)
    c = abi.CoreSource(r_core, 1.0e4, 1.0e43, freq)
    for P in (Pg, Ph):
        P.grid.r_core = r_core
        P.tparams.dt = 3.0
    print("case inner source r_core=%g freq=%g" % (r_core, freq), flush=True)
    assert gpu.emit_inner_source(Pg, c, n, id_base=9, offset=7) == n
    gpu.sync()
    assert OracleBackend().emit_inner_source(Ph, c, n, id_base=9, offset=7) == n
    a, b = Pg.packets_numpy(n + 7), Ph.packets_numpy(n + 7)
    for k in ("ind", "type", "fate", "draw", "id"):
        assert np.array_equal(a[k][7:], b[k][7:]), k
    scale = max(r_core, 1.0)
    for k in ("x", "y", "z"):
        assert np.max(np.abs(a[k][7:] - b[k][7:])) <= 1e-12 * scale, k
    for k in ("Dx", "Dy", "Dz"):
        assert np.max(np.abs(a[k][7:] - b[k][7:])) <= 1e-12, k
    for k in ("t", "e", "nu"):
        assert np.max(np.abs(a[k][7:] - b[k][7:]) / np.abs(b[k][7:])) <= 1e-12, k


@pytest.mark.parametrize("precision", [abi.PREC_MIXED, abi.PREC_FP64_BATCHED])
def test_lightbulb_on_gpu(gpu, precision):
    n = 50_000_000
    P, r_out = lightbulb(runtime.Space("cuda:0"), n)
    P.tparams.precision = precision
    print("case lightbulb precision=%d" % precision, flush=True)
    lightbulb_run(gpu, P, n)
    gpu.sync()
    e_rad, e_gas, e_spec, e_J = check_lightbulb(P, r_out)
    print("T_rad %.5f T_gas %.5f spectrum %.4f J_50 %.4f" % (e_rad, e_gas, e_spec, e_J))
    assert e_rad < 0.01 and e_gas < 0.01
    assert e_spec < 0.10 and e_J < 0.10
