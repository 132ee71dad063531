"""Checkpoint and restart (pubsed_amd.checkpoint; the reference's
writeCheckpointParticles / readCheckpointParticles, transport_output.cpp:
300-365, 501-590, with .npy datasets in place of HDF5): three time-dependent
steps run straight through, and the same run checkpointed after its first
step and restored into a fresh problem, end with the same packets -- every
packet at the same position, direction, energy, frequency, zone, fate and
RNG counter, bit for bit -- and the same tallies and clocks (the tallies up
to the order of the oracle threads' additions).  Radiative equilibrium is
off: it would carry that order into the temperature.  CPU, with the oracle
as the backend."""
import json
import os

import numpy as np
import pytest

from pubsed_amd import runtime, checkpoint, workloads as W
from oracle.oracle_py import OracleBackend

OVER = dict(nz=24, steady=0, radiative_eq=0)
FIELDS = ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu", "ind", "type", "fate", "draw")


def run(P, ob, steps, n_emit, dt):
    for s in steps:
        runtime.time_step(ob, P, dt, n_emit, None, s)


def by_id(P):
    pk = P.packets_numpy(P.n_packets)
    o = np.argsort(pk["id"])
    return {k: v[o] for k, v in pk.items()}


@pytest.mark.parametrize("grid", ["1d", "2d"])
def test_restart_continues_the_run(tmp_path, grid):
    name, over = ("c1_grey", dict(OVER)) if grid == "1d" else ("c4_2d", dict(nx=8, nzz=16, steady=0, radiative_eq=0))
    n_emit, dt = 1500, 0.3 * W.DAY
    ob = OracleBackend()
    A = runtime.Problem(runtime.Space(None), name, **over)
    A.alloc_packets(3 * n_emit)
    run(A, ob, range(3), n_emit, dt)

    B = runtime.Problem(runtime.Space(None), name, **over)
    B.alloc_packets(3 * n_emit)
    run(B, ob, range(1), n_emit, dt)
    meta = checkpoint.write_checkpoint(B, str(tmp_path / "chk"), step_index=0)
    assert meta["n_packets"] == B.n_packets > 0
    assert np.load(str(tmp_path / "chk" / "particles" / "x.npy")).shape == (B.n_packets, 3)

    C = runtime.Problem(runtime.Space(None), name, **over)
    C.alloc_packets(3 * n_emit)
    checkpoint.read_checkpoint(C, str(tmp_path / "chk"))
    run(C, ob, range(1, 3), n_emit, dt)

    assert C.n_packets == A.n_packets and C.t_now == A.t_now and C.grid.t_now == A.grid.t_now
    a, c = by_id(A), by_id(C)
    assert np.array_equal(a["id"], c["id"])
    for k in FIELDS:
        assert np.array_equal(a[k], c[k]), k
    for k in ("rho", "T", "vol"):
        assert np.array_equal(A.get(k), C.get(k)), k
    for k in ("e_abs", "J_nu", "L_radio_dep", "opt_flux", "gam_flux"):
        assert np.allclose(A.get(k), C.get(k), rtol=1e-12, atol=0), k
    # the spectra hold every step's escapes, the first step's from the checkpoint
    for k in ("opt_click", "gam_click"):
        assert np.array_equal(A.get(k), C.get(k)), k


def test_checkpoint_rejects_another_configuration(tmp_path):
    ob = OracleBackend()
    P = runtime.Problem(runtime.Space(None), "c1_grey", **OVER)
    P.alloc_packets(1000)
    runtime.time_step(ob, P, 0.3 * W.DAY, 500, None, 0)
    checkpoint.write_checkpoint(P, str(tmp_path / "chk"))
    Q = runtime.Problem(runtime.Space(None), "c1_grey", **dict(OVER, nz=30))
    Q.alloc_packets(1000)
    with pytest.raises(ValueError):
        checkpoint.read_checkpoint(Q, str(tmp_path / "chk"))
    R = runtime.Problem(runtime.Space(None), "c1_grey", **OVER)
    R.alloc_packets(10)
    with pytest.raises(ValueError):
        checkpoint.read_checkpoint(R, str(tmp_path / "chk"))
    with open(str(tmp_path / "chk" / "meta.json")) as f:
        assert json.load(f)["config"] == "c1_grey"


def test_checkpoint_rejects_another_spectrum_grid(tmp_path):
    ob = OracleBackend()
    P = runtime.Problem(runtime.Space(None), "c1_grey", **OVER)
    P.alloc_packets(1000)
    runtime.time_step(ob, P, 0.3 * W.DAY, 500, None, 0)
    checkpoint.write_checkpoint(P, str(tmp_path / "chk"))
    Q = runtime.Problem(runtime.Space(None), "c1_grey", **dict(OVER, spec_n_mu=4))
    Q.alloc_packets(1000)
    with pytest.raises(ValueError, match="spectrum"):
        checkpoint.read_checkpoint(Q, str(tmp_path / "chk"))
