"""The reference's P-Cygni line (tests/P-Cygni_line) on the GPU: the
problem, criterion and analytic Sobolev profile of test_oracle_pcygni.py
(the line's Sobolev depth 4 from the GPU build's own LTE populations), at
2e7 packets in the mixed form and the batched FP64 form, with the oracle's
bars (mean |MC - analytic| 0.03, largest 0.09 over 1100-1340 A); the GPU
build's line opacity equals the oracle's within 1e-10."""
import numpy as np
import pytest

from pubsed_amd import abi, runtime
from oracle.oracle_py import OracleBackend
from test_oracle_pcygni import pcygni, pcygni_run, check_pcygni, line_strength, sobolev_depth, TAU

pytestmark = pytest.mark.gpu


def test_pcygni_line_opacity_matches_oracle(gpu):
    A_ul = line_strength(OracleBackend(), runtime.Space(None))
    Pg = pcygni(runtime.Space("cuda:0"), 1, A_ul)
    Ph = pcygni(runtime.Space(None), 1, A_ul)
    gpu.opacity_build(Pg)
    gpu.sync()
    OracleBackend().opacity_build(Ph)
    for k in ("abs", "emis_cdf"):
        g, h = Pg.get(k), Ph.get(k)
        assert np.max(np.abs(g - h) / np.maximum(np.abs(h), 1e-300)) < 1e-10, k
    assert abs(sobolev_depth(Pg, A_ul) / TAU - 1) < 1e-10


@pytest.mark.parametrize("precision", [abi.PREC_MIXED, abi.PREC_FP64_BATCHED])
def test_pcygni_on_gpu(gpu, precision):
    n = 20_000_000
    A_ul = line_strength(gpu, runtime.Space("cuda:0"))
    P = pcygni(runtime.Space("cuda:0"), n, A_ul)
    P.tparams.precision = precision
    print("case P-Cygni precision=%d" % precision, flush=True)
    pcygni_run(gpu, P, n)
    gpu.sync()
    assert abs(P.get("opt_click").sum() / n - 1) < 0.5     # most packets escape; the rest hit the core
    mean, worst = check_pcygni(P)
    print("P-Cygni: mean |dF| %.4f, largest %.4f" % (mean, worst))
    assert mean < 0.03 and worst < 0.09
