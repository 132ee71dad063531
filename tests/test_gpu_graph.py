"""A time-dependent step with device-resident packet counts, eager and
captured in a HIP graph per step (sgpu_graph_begin / sgpu_graph_end with
sgpu_emit_radioactive_dev, sgpu_propagate_dev, sgpu_compact_dev), against the
host-count step (runtime.time_step) on the 2D outflow: after every step the
census packets are bit-identical (radiative equilibrium off, so no tally's
summation order reaches the paths), the counts equal, and the tallies and the
accumulated escape spectrum agree to the order of the floating-point atomic
additions (1e-10).  Each captured step updates the previous step's
executable graph in place and is compared with the eager result before the
next one."""
import numpy as np
import pytest
import torch

from pubsed_amd import abi, runtime, workloads as W

pytestmark = pytest.mark.gpu
FIELDS = ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu", "ind", "type", "fate", "draw")


def by_id(P, n):
    pk = P.packets_numpy(n)
    o = np.argsort(pk["id"])
    return {k: v[o] for k, v in pk.items()}


@pytest.mark.parametrize("precision", [abi.PREC_MIXED, abi.PREC_FP64_BATCHED])
def test_graph_step_matches_host_step(gpu, precision):
    over = dict(nx=16, nzz=32, steady=0, radiative_eq=0)
    n_emit, steps, dt = 20000, 4, 0.3 * W.DAY
    cap = n_emit * (steps + 1)

    def problem():
        P = runtime.Problem(runtime.Space("cuda:0"), "c4_2d", **over)
        P.alloc_packets(cap)
        P.tparams.precision = precision
        return P, P.alloc_packets(cap, name="scr")

    A, sa = problem()
    G, sg = problem()
    d_cnt = torch.zeros(2, dtype=torch.int64, device="cuda:0")
    try:
        gpu.use_own_stream()
        for s in range(steps):
            print("case graph step %d precision %d" % (s, precision), flush=True)
            got, alive, esc = runtime.time_step(gpu, A, dt, n_emit, sa, s)
            runtime.time_step_dev(gpu, G, dt, n_emit, sg, s, d_cnt, graph=(s > 0))
            gpu.sync()
            n_g, esc_g = (int(v) for v in d_cnt.cpu())
            if s > 0:
                assert gpu.graph_nodes() >= 10
            assert n_g == alive and esc_g == esc and alive > 0
            a, g = by_id(A, alive), by_id(G, n_g)
            assert np.array_equal(a["id"], g["id"])
            for k in FIELDS:
                assert np.array_equal(a[k], g[k]), (s, k)
            for k in ("e_abs", "J_nu", "L_radio_dep", "opt_flux", "gam_flux"):
                assert np.allclose(A.get(k), G.get(k), rtol=1e-10, atol=0), (s, k)
            for k in ("opt_click", "gam_click", "T", "rho", "vol"):
                assert np.array_equal(A.get(k), G.get(k)), (s, k)
    finally:
        gpu.use_torch_stream()
