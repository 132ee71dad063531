"""Checkpoint and restart on the GPU: three time-dependent steps straight
through, and the same run checkpointed after one step and restored into a
fresh problem, end with bit-identical packets (FP64 drop-in and mixed form:
each packet's path depends only on its own state and RNG counters) and the
same tallies up to the order of the atomic additions."""
import numpy as np
import pytest

from pubsed_amd import runtime, checkpoint, abi, workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu", "ind", "type", "fate", "draw")


def by_id(P):
    pk = P.packets_numpy(P.n_packets)
    o = np.argsort(pk["id"])
    return {k: v[o] for k, v in pk.items()}


@pytest.mark.parametrize("grid,precision", [("1d", abi.PREC_FP64), ("2d", abi.PREC_FP64), ("2d", abi.PREC_MIXED)])
def test_gpu_restart_continues_the_run(gpu, tmp_path, grid, precision):
    name, over = (("c1_grey", dict(nz=24, steady=0, radiative_eq=0)) if grid == "1d"
                  else ("c4_2d", dict(nx=16, nzz=32, steady=0, radiative_eq=0)))
    n_emit, dt = 4000, 0.3 * W.DAY

    def problem():
        P = runtime.Problem(runtime.Space("cuda:0"), name, **over)
        P.alloc_packets(3 * n_emit)
        P.tparams.precision = precision
        return P, P.alloc_packets(3 * n_emit, name="scr")

    print("case checkpoint", grid, precision, flush=True)
    A, sa = problem()
    for s in range(3):
        runtime.time_step(gpu, A, dt, n_emit, sa, s)
    B, sb = problem()
    runtime.time_step(gpu, B, dt, n_emit, sb, 0)
    gpu.sync()
    checkpoint.write_checkpoint(B, str(tmp_path / "chk"), step_index=0)
    C, sc = problem()
    checkpoint.read_checkpoint(C, str(tmp_path / "chk"))
    for s in (1, 2):
        runtime.time_step(gpu, C, dt, n_emit, sc, s)
    gpu.sync()
    print("packets", A.n_packets, C.n_packets)
    assert C.n_packets == A.n_packets > 0 and C.t_now == A.t_now and C.grid.t_now == A.grid.t_now
    a, c = by_id(A), by_id(C)
    assert np.array_equal(a["id"], c["id"])
    for k in FIELDS:
        assert np.array_equal(a[k], c[k]), k
    for k in ("e_abs", "J_nu", "L_radio_dep", "opt_flux", "gam_flux"):
        assert np.allclose(A.get(k), C.get(k), rtol=1e-10, atol=0), k
    # the spectra hold every step's escapes, the first step's from the checkpoint
    for k in ("opt_click", "gam_click"):
        assert np.array_equal(A.get(k), C.get(k)), k
