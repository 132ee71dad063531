"""Checkpoint and restart of a time-dependent run.

The reference checkpoints its particle lists, its random-number streams and
its zone state into HDF5 (transport::writeCheckpointParticles,
transport_output.cpp:300-365, one dataset per particle property -- type, x
(n x 3), D (n x 3), ind, t, e, nu, fate ... -- under the groups
``particles`` and ``particles_escaped``; readCheckpointParticles :501-590).
HDF5 is not available here, so the same groups and datasets are written as
one ``.npy`` file each under a directory: ``<path>/particles/x.npy`` is the
reference's ``particles/x`` dataset.  The random-number state of a packet
is its Philox stream (seed, id) and the index of its next unused block
(``draw``), so ``particles/draw`` and ``particles/id`` are the RNG
checkpoint.  Escaped packets are not kept (the compaction drops them after
the spectrum has counted them), so there is no ``particles_escaped`` group.

Besides the particles: the zone state that the next step reads (density,
temperature, grid edges and volumes, which the homologous step has moved),
the radiation tallies (J_nu feeds the next step's NLTE gas state), the
optical and gamma escape spectra (flux and click counts, which a
time-dependent run keeps adding to), the clocks and the grid's inner edges,
in ``state/*.npy`` and ``meta.json``.
A run restored into a Problem of the same configuration continues as the
uninterrupted run does: the same packets take the same paths, and the
tallies differ only by the order of their floating-point atomic additions
(under radiative equilibrium that order reaches the temperature, and so the
packets, at the level of rounding).
"""
import json
import os

import numpy as np

PARTICLE_SCALARS = ("type", "ind", "t", "e", "nu", "fate", "draw", "id")
TALLIES = ("e_abs", "J_nu", "fx_rad", "fy_rad", "fz_rad", "fr_rad", "L_radio_dep", "L_radio_emit", "e_rad",
           # the escape spectra, which accumulate over the steps of a
           # time-dependent run (transport::writeCheckpointSpectra,
           # transport_output.cpp:492, read back at transport_init.cpp:164)
           "opt_flux", "opt_click", "gam_flux", "gam_click")
FORMAT = "sedona-mi355x checkpoint 1"


def _zone_keys(P):
    grid = ("r_out",) if "r_out" in P.a else ("x_out", "z_out")
    return ("rho", "T", "vol") + grid


def write_checkpoint(P, path, step_index=None):
    """Write the packets on the grid, the zone state, the tallies and the
    clocks of Problem P under the directory ``path``."""
    n = P.n_packets
    pk = P.packets_numpy(n)
    os.makedirs(os.path.join(path, "particles"), exist_ok=True)
    os.makedirs(os.path.join(path, "state"), exist_ok=True)
    sv = lambda sub, k, v: np.save(os.path.join(path, sub, k + ".npy"), np.ascontiguousarray(v),
                                   allow_pickle=False)
    sv("particles", "x", np.stack([pk["x"], pk["y"], pk["z"]], axis=1))
    sv("particles", "D", np.stack([pk["Dx"], pk["Dy"], pk["Dz"]], axis=1))
    for k in PARTICLE_SCALARS:
        sv("particles", k, pk[k])
    for k in _zone_keys(P) + TALLIES:
        sv("state", k, P.get(k))
    meta = {"format": FORMAT, "config": P.name, "seed": P.seed, "n_zones": P.nz, "n_nu": P.n_nu,
            "n_packets": n, "t_now": P.t_now, "grid_t_now": P.grid.t_now,
            "grid_min": [P.grid.r_out.min, P.grid.x_out.min, P.grid.z_out.min],
            "first_step": int(P.op_params.first_step), "rng_seed": int(P.tparams.seed),
            "step_index": step_index}
    with open(os.path.join(path, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    return meta


def read_checkpoint(P, path):
    """Restore a checkpoint into Problem P, which must have the checkpoint's
    configuration (zone and frequency counts) and room for its packets.
    Returns the checkpoint's metadata."""
    with open(os.path.join(path, "meta.json")) as f:
        meta = json.load(f)
    if meta.get("format") != FORMAT:
        raise ValueError("%s: not a checkpoint of this format" % path)
    if meta["n_zones"] != P.nz or meta["n_nu"] != P.n_nu:
        raise ValueError("checkpoint of %d zones x %d bins, problem of %d x %d"
                         % (meta["n_zones"], meta["n_nu"], P.nz, P.n_nu))
    n = int(meta["n_packets"])
    if n and (P.packets is None or n > P.packets.capacity):
        raise ValueError("checkpoint holds %d packets; the problem has room for %d"
                         % (n, 0 if P.packets is None else P.packets.capacity))
    ld = lambda sub, k: np.load(os.path.join(path, sub, k + ".npy"), allow_pickle=False)
    state = {k: ld("state", k) for k in _zone_keys(P) + TALLIES}
    for k, v in state.items():
        have = P.a[k].numel() if P.space.on_gpu else P.a[k].size
        if v.size != have:
            raise ValueError("checkpoint dataset state/%s holds %d values, the problem %d (another "
                             "spectrum grid?)" % (k, v.size, have))
    for k, v in state.items():
        P.set(k, v)
    if n:
        x, D = ld("particles", "x"), ld("particles", "D")
        arrays = {"x": x[:, 0], "y": x[:, 1], "z": x[:, 2], "Dx": D[:, 0], "Dy": D[:, 1], "Dz": D[:, 2]}
        for k in PARTICLE_SCALARS:
            arrays[k] = ld("particles", k)
        for k, v in arrays.items():
            dst = P.a["pk_" + k]
            v = np.ascontiguousarray(v)
            if v.dtype in (np.uint32, np.uint64):
                v = v.view(np.int32 if v.dtype == np.uint32 else np.int64)
            if P.space.on_gpu:
                dst[:n].copy_(P.space.torch.from_numpy(v).to(P.space.device))
            else:
                dst[:n] = v.view(dst.dtype) if v.dtype != dst.dtype else v
    P.n_packets = n
    P.t_now = meta["t_now"]
    P.grid.t_now = meta["grid_t_now"]
    P.grid.r_out.min, P.grid.x_out.min, P.grid.z_out.min = meta["grid_min"]
    P.tparams.t_now = P.op_params.t_now = P.t_now
    P.op_params.first_step = meta["first_step"]
    P.tparams.seed = meta["rng_seed"]
    return meta
