"""Device plumbing for the transport hot path: where arrays live, how the
C-ABI descriptors are filled, and the named workloads.

A Problem holds every input / output array of one transport configuration
in one Space (host numpy arrays for the CPU oracle, or torch tensors on a
HIP device for libsedona_gpu.so) together with the ctypes descriptors that
point at them.  Backends (GpuBackend here; the CPU oracle's wrapper lives
with the oracle) take a Problem and run the reference's step functions.
"""
import ctypes
import math

import numpy as np

from . import abi, atomdata, workloads as W

DTYPES = {"f64": np.float64, "f32": np.float32, "i32": np.int32, "u32": np.uint32,
          "i64": np.int64, "u64": np.uint64}


class Space:
    """Host (device=None: numpy) or a torch HIP device ('cuda:0')."""

    def __init__(self, device=None):
        self.device = device
        self.keep = []
        if device is not None:
            import torch
            self.torch = torch

    @property
    def on_gpu(self):
        return self.device is not None

    def _torch_dtype(self, dt):
        t = self.torch
        return {np.dtype(np.float64): t.float64, np.dtype(np.float32): t.float32,
                np.dtype(np.int32): t.int32, np.dtype(np.uint32): t.int32,
                np.dtype(np.int64): t.int64, np.dtype(np.uint64): t.int64}[np.dtype(dt)]

    def asarray(self, a, dtype=None):
        a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
        if not self.on_gpu:
            a = a.copy()
            self.keep.append(a)
            return a
        t = self.torch
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        elif a.dtype == np.uint64:
            a = a.view(np.int64)
        x = t.from_numpy(a.copy()).to(self.device)
        self.keep.append(x)
        return x

    def zeros(self, n, dtype):
        dt = DTYPES.get(dtype, dtype)
        if not self.on_gpu:
            a = np.zeros(n, dtype=dt)
            self.keep.append(a)
            return a
        x = self.torch.zeros(n, dtype=self._torch_dtype(dt), device=self.device)
        self.keep.append(x)
        return x

    def to_numpy(self, x, dtype=None):
        if x is None:
            return None
        if not self.on_gpu:
            a = np.array(x)
        else:
            a = x.detach().cpu().numpy()
        if dtype is not None:
            a = a.view(DTYPES.get(dtype, dtype))
        return a

    def sync(self):
        if self.on_gpu:
            self.torch.cuda.synchronize()


def locate_struct(space, g):
    x = space.asarray(np.asarray(g["x"], np.float64))
    return abi.Locate(int(g["n"]), int(g["type"]), float(g["min"]), float(g["del_"]), abi.ptr(x)), x


# ----------------------------------------------------------------------
# named configurations (sizes and physics of each transport workload)
# ----------------------------------------------------------------------
C1_ELEMS = ((28, 56), (27, 56), (26, 56), (14, 28))
KAPPA_ABS_C1 = 0.1
KAPPA_T = W.THOMSON * 0.5 / W.M_P   # Thomson opacity for Y_e = 0.5 (cm^2/g)

CONFIGS = {
    # grey 1D: kappa_abs 0.1 + Thomson as one grey opacity with
    # epsilon = kappa_abs / kappa_tot (grey mode has no separate e-scattering)
    "c1_grey": dict(grid="1d", nz=100, nu=(1e14, 1e16, 1024), elements=C1_ELEMS,
                    grey=KAPPA_ABS_C1 + KAPPA_T, epsilon=KAPPA_ABS_C1 / (KAPPA_ABS_C1 + KAPPA_T),
                    escat=0, fuzz=0, radiative_eq=0, steady=1, n_packets=1_000_000),
    # line-expansion 1D: e-scattering + fuzz-line expansion opacity, LTE
    # Saha-Boltzmann for 4 ion stages of each element, epsilon = 1 and
    # radiative equilibrium (every absorption re-emits), as the reference's
    # toy SN Ia spectrum test (tests/toy_type1a_supernova/1D_spectrum/param.lua)
    "c2_line": dict(grid="1d", nz=128, nu=(W.C_LIGHT / 1e-4, W.C_LIGHT / 1e-5, 4096),
                    elements=((28, 56), (14, 28)), grey=0.0, epsilon=1.0, escat=1, fuzz=1,
                    radiative_eq=1, steady=1, n_lines=100_000, n_packets=10_000_000),
    # 2D (varpi, z) 128 x 256 axis ratio 2, time-dependent with census; the
    # escape spectrum in 10 mu bins and 0.1-day time bins as the reference's
    # 2D benchmark (tests/ellipsoidal_outflow_2D/2D_benchmark/param.lua:45-48),
    # the time grid starting 2 d before the ejecta's 20 d (an escaping
    # packet's t_obs = t - x.D / c reaches back v_max t / c = 1.3 d)
    "c4_2d": dict(grid="2d", nx=128, nzz=256, nu=(W.C_LIGHT / 1e-4, W.C_LIGHT / 1e-5, 4096),
                  elements=((28, 56), (14, 28)), grey=0.0, epsilon=1.0, escat=1, fuzz=1,
                  radiative_eq=1, steady=0, n_lines=100_000, n_packets=100_000_000,
                  spec_t0=18 * W.DAY, spec_t1=28 * W.DAY, spec_dt=0.1 * W.DAY, spec_n_mu=10),
}


class Problem:
    """Every array of one transport configuration in one Space."""

    def __init__(self, space, name="c1_grey", seed=20261203, **over):
        cfg = dict(CONFIGS[name])
        cfg.update(over)
        self.name, self.cfg, self.space, self.seed = name, cfg, space, seed
        S = space
        a = self.a = {}
        lo, hi, nbin = cfg["nu"]
        # transport_nu_grid {nu_1, nu_2, del} with nu_1 == nu_2 (the default
        # {1,1,1}) is locate_array's catch-all single bin: grey transport
        # {nu_1, nu_2, del, 1} with a given del: locate_array::log_init's own edges
        if lo == hi:
            self.nu_np = W.lin_init_n(lo, hi, 1)
        elif cfg.get("nu_log_del"):
            x = W.log_init(lo, hi, cfg["nu_log_del"])
            self.nu_np = dict(n=len(x), type=abi.LOC_LOG, min=lo, del_=cfg["nu_log_del"], x=x)
        else:
            self.nu_np = W.nu_log_grid(lo, hi, nbin)
        self.nu, a["nu_x"] = locate_struct(S, self.nu_np)
        self.n_nu = nn = self.nu_np["n"]
        if cfg["grid"] == "1d":
            ej = W.ejecta_1d(nz=cfg["nz"], elements=cfg["elements"])
            self.ej = ej
            nz = ej["nz"]
            a["r_out"] = S.asarray(ej["r_out"])
            a["vol"] = S.asarray(ej["vol"])
            r_loc = abi.Locate(nz, abi.LOC_FLEX, 0.0, 0.0, abi.ptr(a["r_out"]))
            self.grid = abi.Grid(abi.GRID_1D_SPHERE, nz, ej["t"], 0.0, r_loc, abi.Locate(),
                                 abi.Locate(), abi.ptr(a["vol"]))
        else:
            # a model read from a file (models.read_model_2d), or the synthetic ejecta
            ej = cfg["model"] if cfg.get("model") is not None else \
                W.ejecta_2d(nx=cfg["nx"], nzz=cfg["nzz"], elements=cfg["elements"])
            self.ej = ej
            nz = ej["nz"]
            a["x_out"] = S.asarray(ej["x_out"])
            a["z_out"] = S.asarray(ej["z_out"])
            a["vol"] = S.asarray(ej["vol"])
            xl = abi.Locate(ej["nx"], abi.LOC_FLEX, ej["x_min"], 0.0, abi.ptr(a["x_out"]))
            zl = abi.Locate(ej["nzz"], abi.LOC_FLEX, ej["z_min"], 0.0, abi.ptr(a["z_out"]))
            self.grid = abi.Grid(abi.GRID_2D_CYLN, nz, ej["t"], 0.0, abi.Locate(), xl, zl,
                                 abi.ptr(a["vol"]))
        self.nz = nz
        ne = len(ej["elem_Z"])
        a["rho"] = S.asarray(ej["rho"])
        a["T"] = S.asarray(ej["T"])
        a["X"] = S.asarray(ej["X"].ravel())
        a["grey"] = S.asarray(np.full(nz, cfg["grey"]))
        a["elem_Z"] = S.asarray(ej["elem_Z"])
        a["elem_A"] = S.asarray(ej["elem_A"])
        self.zs = abi.ZoneState(nz, ne, abi.ptr(a["rho"]), abi.ptr(a["T"]), abi.ptr(a["X"]), None,
                                abi.ptr(a["grey"]), abi.ptr(a["elem_Z"]), abi.ptr(a["elem_A"]))
        self.atoms = None
        self.n_ion = self.n_lev = 0
        if cfg["grey"] == 0:
            if cfg.get("atom_file"):             # a Sedona atomic data file (data/*.hdf5)
                at = atomdata.read_atomic_data(cfg["atom_file"], ej["elem_Z"], self.nu_np,
                                               cfg.get("max_ion_stage", 9999))
            elif cfg.get("atom") is not None:    # one hand-built atom (hydrogen_atom's arguments)
                assert len(ej["elem_Z"]) == 1
                at = W.hydrogen_atom(self.nu_np, **cfg["atom"])
            else:
                at = W.synthetic_atoms(self.nu_np, ej["elem_Z"], n_lines=cfg.get("n_lines", 100000),
                                       seed=seed)
            self.atoms_np = at
            for k in ("lev_off", "ion_off", "lev_E", "lev_g", "lev_ion", "ion_chi", "atom_zero_eps",
                      "line_nu", "line_gf", "line_El", "line_ion", "line_bin", "line_bin_off",
                      "bline_l", "bline_u", "bline_nu", "bline_f", "bline_A", "bline_Bul",
                      "bline_Blu", "bline_bin", "bline_bin_off"):
                a["at_" + k] = S.asarray(at[k])
            self.n_lev = int(at["lev_off"][-1])
            self.n_ion = int(at["ion_off"][-1])
            self.atoms = abi.Atoms(at["n_atoms"], at["max_levels"], self.n_lev, self.n_ion,
                                   *[abi.ptr(a["at_" + k]) for k in
                                     ("lev_off", "ion_off", "lev_E", "lev_g", "lev_ion", "ion_chi",
                                      "atom_zero_eps")],
                                   len(at["line_nu"]),
                                   *[abi.ptr(a["at_" + k]) for k in
                                     ("line_nu", "line_gf", "line_El", "line_ion", "line_bin",
                                      "line_bin_off")],
                                   len(at["bline_nu"]),
                                   *[abi.ptr(a["at_" + k]) for k in
                                     ("bline_l", "bline_u", "bline_nu", "bline_f", "bline_A",
                                      "bline_Bul", "bline_Blu", "bline_bin", "bline_bin_off")])
        self.t_now = ej["t"]
        self.op_params = abi.OpacityParams(
            self.t_now, 1.0 if cfg["steady"] else cfg.get("dt", 0.1 * W.DAY), cfg["epsilon"],
            1000.0, 1e8, 1e40, 0.0, 0.0, cfg["escat"], cfg["fuzz"], 0, 0, cfg["radiative_eq"], 0,
            cfg.get("lexp", 0), cfg.get("ff", 0), cfg.get("bf", 0), cfg.get("no_ground_recomb", 0),
            cfg.get("bb", 0))
        self.op_params.line_velocity_width = cfg.get("line_velocity_width", 0.0)
        self.op_params.minimum_extinction = cfg.get("minimum_extinction", 0.0)
        for k in ("abs", "scat", "emis_cdf"):
            a[k] = S.zeros(nz * nn, "f64")
        for k in ("L_thermal", "planck_mean", "rosseland_mean", "n_elec", "compton_opac",
                  "photoion_opac", "eps_imc"):
            a[k] = S.zeros(nz, "f64")
        a["zone_status"] = S.zeros(nz, "i32")
        a["ion_frac"] = S.zeros(max(nz * self.n_ion, 1), "f64")
        a["ion_part"] = S.zeros(max(nz * self.n_ion, 1), "f64")
        a["lev_n"] = S.zeros(max(nz * self.n_lev, 1), "f64")
        P = abi.ptr
        self.op_out = abi.OpacityOut(P(a["abs"]), P(a["scat"]), P(a["emis_cdf"]), P(a["L_thermal"]),
                                     P(a["planck_mean"]), P(a["rosseland_mean"]), P(a["n_elec"]),
                                     P(a["compton_opac"]), P(a["photoion_opac"]), P(a["eps_imc"]),
                                     P(a["zone_status"]),
                                     P(a["ion_frac"]) if self.n_ion else None,
                                     P(a["ion_part"]) if self.n_ion else None,
                                     P(a["lev_n"]) if self.n_lev else None)
        self.tables = abi.Tables(P(a["abs"]), P(a["scat"]), P(a["emis_cdf"]), P(a["eps_imc"]),
                                 P(a["compton_opac"]), P(a["photoion_opac"]))
        steady = cfg["steady"]
        self.tparams = abi.TransportParams(self.t_now, 1.0 if steady else cfg.get("dt", 0.1 * W.DAY),
                                           seed, steady, cfg["radiative_eq"], 0, 0, 1, 0,
                                           abi.PREC_FP64, 0)
        # tallies and spectra (spectrum_time_grid default {0,1,1}: one bin;
        # spectrum_n_mu / spectrum_n_phi default 1; the gamma spectrum uses
        # the optical grids, as transport_init.cpp:166-173 does; mu on
        # [-1, 1] and phi on [0, 2 pi] in equal bins, spectrum_array.cpp:66-70)
        for k in ("e_abs", "fx_rad", "fy_rad", "fz_rad", "fr_rad", "L_radio_dep",
                  "L_radio_emit", "e_rad"):
            a[k] = S.zeros(nz, "f64")
        a["J_nu"] = S.zeros(nz * nn, "f64")
        tg = W.lin_init_del(0.0, 1.0, 1.0) if steady else W.lin_init_del(
            cfg.get("spec_t0", 0.0), cfg.get("spec_t1", 1e8), cfg.get("spec_dt", 1e8))
        n_mu, n_phi = int(cfg.get("spec_n_mu", 1)), int(cfg.get("spec_n_phi", 1))
        self.spec_grids = dict(time=tg, nu=self.nu_np, mu=W.lin_init_n(-1, 1, n_mu),
                               phi=W.lin_init_n(0, 2 * W.PI_REF, n_phi))
        self.spec_shape = (tg["n"], nn, self.spec_grids["mu"]["n"], self.spec_grids["phi"]["n"])
        self.n_spec = int(np.prod(self.spec_shape))
        specs = []
        for nm in ("opt", "gam"):
            a[nm + "_flux"] = S.zeros(self.n_spec, "f64")
            a[nm + "_click"] = S.zeros(self.n_spec, "f64")
            locs = []
            for gk in ("time", "nu", "mu", "phi"):
                l, x = locate_struct(S, self.spec_grids[gk])
                a["%s_%s_x" % (nm, gk)] = x
                locs.append(l)
            specs.append(abi.Spectrum(*locs, P(a[nm + "_flux"]), P(a[nm + "_click"])))
        self.tallies = abi.Tallies(*[P(a[k]) for k in ("e_abs", "J_nu", "fx_rad", "fy_rad", "fz_rad",
                                                         "fr_rad", "L_radio_dep", "L_radio_emit",
                                                         "e_rad")], specs[0], specs[1])
        self.packets = None
        self.n_packets = 0

    def alloc_packets(self, capacity, name="pk"):
        S, a = self.space, self.a
        for k in ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu"):
            a["%s_%s" % (name, k)] = S.zeros(capacity, "f64")
        for k in ("ind", "type", "fate"):
            a["%s_%s" % (name, k)] = S.zeros(capacity, "i32")
        a[name + "_draw"] = S.zeros(capacity, "u32")
        a[name + "_id"] = S.zeros(capacity, "u64")
        pk = abi.Packets(capacity, *[abi.ptr(a["%s_%s" % (name, k)]) for k in
                                     ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu", "ind", "type",
                                      "fate", "draw", "id")])
        if name == "pk":
            self.packets = pk
        return pk

    def packets_numpy(self, n=None, name="pk"):
        a, S = self.a, self.space
        out = {}
        for k in ("x", "y", "z", "Dx", "Dy", "Dz", "t", "e", "nu", "ind", "type", "fate"):
            out[k] = S.to_numpy(a["%s_%s" % (name, k)])
        out["draw"] = S.to_numpy(a[name + "_draw"]).view(np.uint32)
        out["id"] = S.to_numpy(a[name + "_id"]).view(np.uint64)
        if n is not None:
            out = {k: v[:n] for k, v in out.items()}
        return out

    def load_packets(self, arrs, name="pk"):
        """Copy host packet arrays (dict of numpy) into the packet buffers."""
        a, S = self.a, self.space
        n = len(arrs["x"])
        for k, v in arrs.items():
            dst = a["%s_%s" % (name, k)]
            v = np.ascontiguousarray(v)
            if v.dtype == np.uint32:
                v = v.view(np.int32)
            elif v.dtype == np.uint64:
                v = v.view(np.int64)
            if S.on_gpu:
                dst[:n].copy_(S.torch.from_numpy(v).to(S.device))
            else:
                dst[:n] = v
        return n

    def get(self, key, dtype=None):
        return self.space.to_numpy(self.a[key], dtype)

    def set(self, key, values):
        dst = self.a[key]
        v = np.ascontiguousarray(values)
        if self.space.on_gpu:
            dst.copy_(self.space.torch.from_numpy(v).to(self.space.device))
        else:
            dst[...] = v


def time_step(backend, P, dt, n_emit, scratch_pk, step_index, counters=None):
    """One time step of a time-dependent run in the reference's order
    (SedonaClass::evolve_system, SedonaClass.cpp:346-355, then
    transport::step, transport.cpp:26-140): the homologous hydro step moves
    the grid to t + dt first, then opacities, wipe, emission appended after
    the census packets, propagation to t + dt, compaction (census packets
    persist), normalisation and, in radiative equilibrium, the gas
    temperature for the next step; the transport clock then advances by dt.
    Returns (packets emitted, packets alive after the step, escaped)."""
    backend.homologous_step(P, dt)
    P.tparams.dt = dt
    P.op_params.dt = dt
    if step_index == 0:
        P.op_params.first_step = 1     # transport_init.cpp:251, cleared at transport.cpp:138
    nlte = P.cfg.get("nlte_atoms")     # opacity_atoms_in_nlte: per-atom flags
    if nlte is not None and P.atoms is not None:
        # transport_opacity.cpp:109-137: the gas state from the previous
        # step's J_nu (LTE on the first step, inside the call)
        backend.opacity_build_nlte(P, P.a["J_nu"], nlte, P.cfg.get("use_collisions_nlte", 1))
    else:
        backend.opacity_build(P)
    # transport::wipe_radiation clears the zone tallies every step; the
    # escape spectra only in steady iteration (SedonaClass.cpp:349), so a
    # time-dependent run's spectrum holds every step's escapes
    backend.wipe(P, wipe_spectra=int(P.cfg["steady"]))
    n0 = P.n_packets
    got = backend.emit(P, n_emit, id_base=(step_index + 1) << 32, offset=n0)
    n = n0 + got
    backend.propagate(P, n, counters)
    alive, esc = backend.compact(P, n, scratch_pk)
    P.n_packets = alive
    backend.normalize(P)
    if P.cfg["radiative_eq"]:
        backend.solve_eq_temperature(P)   # transport.cpp:126-131
    P.t_now += dt
    P.tparams.t_now = P.t_now
    P.op_params.t_now = P.t_now
    P.op_params.first_step = 0
    return got, alive, esc


def time_step_dev(backend, P, dt, n_emit, scratch_pk, step_index, d_cnt, counters=None, graph=False):
    """time_step with the packet counts kept on the device -- d_cnt, a device
    int64 pair: [0] packets on the grid (in and out), [1] escapes of the step
    -- so that the step is enqueued without a host round trip; with
    graph=True it is captured into a HIP graph and launched (the backend must
    run on its own stream, and a run's first step uncaptured, so that every
    call's scratch exists).  LTE opacities only (the NLTE gas state iterates
    with the host)."""
    backend.homologous_step(P, dt)
    P.tparams.dt = dt
    P.op_params.dt = dt
    P.op_params.first_step = 1 if step_index == 0 else 0
    if graph:
        backend.graph_begin()
    try:
        backend.opacity_build(P, zone_status=False)
        backend.wipe(P, wipe_spectra=int(P.cfg["steady"]))
        backend.emit_dev(P, n_emit, d_cnt, id_base=(step_index + 1) << 32)
        backend.propagate_dev(P, d_cnt, counters)
        backend.compact_dev(P, scratch_pk, d_cnt, d_cnt[1:])
        backend.normalize(P)
        if P.cfg["radiative_eq"]:
            backend.solve_eq_temperature(P)
    except Exception:
        if graph:
            try:
                backend.graph_end(launch=False)
            except Exception:
                pass
        raise
    if graph:
        backend.graph_end(launch=True)
    P.t_now += dt
    P.tparams.t_now = P.t_now
    P.op_params.t_now = P.t_now
    P.op_params.first_step = 0


class GpuBackend:
    """Runs the reference's step functions through libsedona_gpu.so."""

    def __init__(self, device=0, stream=None):
        from . import load_library, SgpuError
        self.lib = load_library()
        self.ctx = self.lib.sgpu_create(device)
        if not self.ctx:
            from . import last_error
            raise SgpuError("sgpu_create failed: " + last_error())
        if stream is not None:
            self.lib.sgpu_set_stream(self.ctx, stream)

    def close(self):
        if self.ctx:
            self.lib.sgpu_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_torch_stream(self):
        import torch
        self.lib.sgpu_set_stream(self.ctx, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))

    def sync(self):
        from . import check
        check(self.lib.sgpu_synchronize(self.ctx), "sgpu_synchronize")

    def opacity_build(self, P, zone_status=True):
        """zone_status=False leaves the per-zone solve status unwritten, so
        the call makes no host round trip (it can then be graph-captured)."""
        from . import check
        out = P.op_out
        if not zone_status:
            out = abi.OpacityOut(*[getattr(P.op_out, f) for f, _ in abi.OpacityOut._fields_])
            out.zone_status = None
        return check(self.lib.sgpu_opacity_build(self.ctx, ctypes.byref(P.nu), ctypes.byref(P.zs),
                                                 ctypes.byref(P.atoms) if P.atoms else None,
                                                 ctypes.byref(P.op_params), ctypes.byref(out)),
                     "sgpu_opacity_build")

    # device-resident packet counts and HIP graphs of a step
    def use_own_stream(self):
        from . import check
        check(self.lib.sgpu_use_own_stream(self.ctx), "sgpu_use_own_stream")

    def graph_begin(self):
        from . import check
        check(self.lib.sgpu_graph_begin(self.ctx), "sgpu_graph_begin")

    def graph_end(self, launch=True):
        from . import check
        check(self.lib.sgpu_graph_end(self.ctx, 1 if launch else 0), "sgpu_graph_end")

    def graph_nodes(self):
        return int(self.lib.sgpu_graph_nodes(self.ctx))

    def emit_dev(self, P, n_emit, d_count, id_base=0, force_rproc=0):
        from . import check
        check(self.lib.sgpu_emit_radioactive_dev(self.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu),
                                                 ctypes.byref(P.zs), ctypes.byref(P.tables),
                                                 ctypes.byref(P.tparams), n_emit, force_rproc, id_base,
                                                 ctypes.byref(P.packets), abi.ptr(d_count),
                                                 ctypes.byref(P.tallies)), "sgpu_emit_radioactive_dev")

    def propagate_dev(self, P, d_n, counters=None):
        from . import check
        check(self.lib.sgpu_propagate_dev(self.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu),
                                          ctypes.byref(P.tables), ctypes.byref(P.tparams),
                                          ctypes.byref(P.packets), abi.ptr(d_n), ctypes.byref(P.tallies),
                                          abi.ptr(counters)), "sgpu_propagate_dev")

    def compact_dev(self, P, scratch_pk, d_count, d_escaped=None):
        from . import check
        check(self.lib.sgpu_compact_dev(self.ctx, ctypes.byref(P.packets), abi.ptr(d_count),
                                        ctypes.byref(scratch_pk), abi.ptr(d_escaped)), "sgpu_compact_dev")

    def emit(self, P, n_emit, id_base=0, offset=0, force_rproc=0):
        from . import check
        nout = ctypes.c_int64(0)
        check(self.lib.sgpu_emit_radioactive(self.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu),
                                             ctypes.byref(P.zs), ctypes.byref(P.tables),
                                             ctypes.byref(P.tparams), n_emit, force_rproc, id_base,
                                             ctypes.byref(P.packets), offset, ctypes.byref(P.tallies),
                                             ctypes.byref(nout)), "sgpu_emit_radioactive")
        return nout.value

    def emit_inner_source(self, P, core, n_emit, id_base=0, offset=0):
        """sgpu_emit_inner_source with core an abi.CoreSource."""
        from . import check
        nout = ctypes.c_int64(0)
        check(self.lib.sgpu_emit_inner_source(self.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu),
                                              ctypes.byref(core), ctypes.byref(P.tparams), n_emit, id_base,
                                              ctypes.byref(P.packets), offset, ctypes.byref(nout)),
              "sgpu_emit_inner_source")
        return nout.value

    def initialize_particles(self, P, n_init, id_base=0, offset=0):
        """sgpu_initialize_particles from P's e_rad tally; rewrites P's
        emissivity CDF with the zones' blackbody distributions."""
        from . import check
        nout = ctypes.c_int64(0)
        check(self.lib.sgpu_initialize_particles(self.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu),
                                                 ctypes.byref(P.zs), abi.ptr(P.a["e_rad"]),
                                                 abi.ptr(P.a["emis_cdf"]), ctypes.byref(P.tparams), n_init,
                                                 id_base, ctypes.byref(P.packets), offset, ctypes.byref(nout)),
              "sgpu_initialize_particles")
        return nout.value

    def opacity_build_nlte(self, P, J_nu, atom_nlte, use_collisions=1, lu_precision=None):
        """sgpu_opacity_build_nlte: J_nu a device tensor [nz][n_nu], atom_nlte a
        host sequence of per-atom flags.  Returns (status, n_e evaluation passes)."""
        from . import check
        import numpy as np
        fl = np.ascontiguousarray(atom_nlte, dtype=np.int32)
        ev = ctypes.c_int32(0)
        prec = abi.NLTE_FP64 if lu_precision is None else lu_precision
        st = check(self.lib.sgpu_opacity_build_nlte(self.ctx, ctypes.byref(P.nu), ctypes.byref(P.zs),
                                                     ctypes.byref(P.atoms), ctypes.byref(P.op_params),
                                                     abi.ptr(J_nu), fl.ctypes.data_as(ctypes.c_void_p),
                                                     use_collisions, prec,
                                                     ctypes.byref(P.op_out), ctypes.byref(ev)),
                   "sgpu_opacity_build_nlte")
        return st, ev.value

    def propagate(self, P, n, counters=None):
        """counters: a torch int64 tensor of 10 entries on the device (or None)."""
        from . import check
        return check(self.lib.sgpu_propagate(self.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu),
                                             ctypes.byref(P.tables), ctypes.byref(P.tparams),
                                             ctypes.byref(P.packets), n, ctypes.byref(P.tallies),
                                             abi.ptr(counters)), "sgpu_propagate")

    def compact(self, P, n, scratch_pk):
        from . import check
        na, ne = ctypes.c_int64(0), ctypes.c_int64(0)
        check(self.lib.sgpu_compact(self.ctx, ctypes.byref(P.packets), n, ctypes.byref(scratch_pk),
                                    ctypes.byref(na), ctypes.byref(ne)), "sgpu_compact")
        return na.value, ne.value

    def wipe(self, P, wipe_spectra=1):
        from . import check
        return check(self.lib.sgpu_wipe_radiation(self.ctx, P.nz, P.n_nu, ctypes.byref(P.tallies),
                                                  wipe_spectra), "sgpu_wipe_radiation")

    def normalize(self, P):
        from . import check
        return check(self.lib.sgpu_normalize_tallies(self.ctx, ctypes.byref(P.grid), ctypes.byref(P.nu),
                                                     ctypes.byref(P.tparams), ctypes.byref(P.tallies)),
                     "sgpu_normalize_tallies")

    def solve_eq_temperature(self, P, set_Tgas_to_Trad=0, n_eval=None):
        """transport::solve_eq_temperature: P's T array from its e_abs tally."""
        from . import check
        return check(self.lib.sgpu_solve_eq_temperature(
            self.ctx, ctypes.byref(P.nu), P.nz, abi.ptr(P.a["abs"]), abi.ptr(P.a["e_abs"]),
            abi.ptr(P.a["e_rad"]), P.op_params.temp_min, P.op_params.temp_max, set_Tgas_to_Trad,
            abi.ptr(P.a["T"]), abi.ptr(n_eval)), "sgpu_solve_eq_temperature")

    def homologous_step(self, P, dt):
        """hydro_homologous::step: rho /= e^3, edges and volumes scaled by
        e = (t+dt)/t, the grid's clock advanced by dt."""
        from . import check
        e = (P.grid.t_now + dt) / P.grid.t_now
        a = P.a
        check(self.lib.sgpu_homologous_step(self.ctx, ctypes.byref(P.grid), abi.ptr(a["rho"]),
                                            abi.ptr(a.get("r_out")), abi.ptr(a.get("x_out")),
                                            abi.ptr(a.get("z_out")), abi.ptr(a["vol"]), e),
              "sgpu_homologous_step")
        P.grid.t_now = P.grid.t_now + dt

    def nlte_solve(self, n, rates, lte, b, pop, work):
        from . import check
        return check(self.lib.sgpu_nlte_solve(self.ctx, n, abi.ptr(rates), abi.ptr(lte), abi.ptr(b),
                                              abi.ptr(pop), abi.ptr(work)), "sgpu_nlte_solve")

    def nlte_batched(self, nsys, n, rate_off, vec_off, rates, lte, b, pop, precision, tol, iters):
        from . import check
        return check(self.lib.sgpu_nlte_batched(self.ctx, nsys, abi.ptr(n), abi.ptr(rate_off),
                                                abi.ptr(vec_off), abi.ptr(rates), abi.ptr(lte),
                                                abi.ptr(b), abi.ptr(pop), precision, tol,
                                                abi.ptr(iters)), "sgpu_nlte_batched")

    def atomic_rate(self, buf, adds):
        """Kernel time (ms) of `adds` FP64 atomic additions of 1.0 at
        hash-scattered indices of the float64 device tensor buf."""
        from . import check
        ms = ctypes.c_double(0)
        check(self.lib.sgpu_atomic_rate(self.ctx, abi.ptr(buf), buf.numel(), adds, ctypes.byref(ms)),
              "sgpu_atomic_rate")
        return ms.value

    def nlte_set_tuning(self, precision, threads, panel, wg_per_cu):
        from . import check
        return check(self.lib.sgpu_nlte_set_tuning(self.ctx, precision, threads, panel, wg_per_cu),
                     "sgpu_nlte_set_tuning")

    def nlte_autotune(self, nsys, n, rate_off, vec_off, rates, lte, b, pop, iters):
        from . import check
        buf = ctypes.create_string_buffer(4096)
        check(self.lib.sgpu_nlte_autotune(self.ctx, nsys, abi.ptr(n), abi.ptr(rate_off),
                                          abi.ptr(vec_off), abi.ptr(rates), abi.ptr(lte), abi.ptr(b),
                                          abi.ptr(pop), abi.ptr(iters), buf, 4096),
              "sgpu_nlte_autotune")
        return buf.value.decode()

    def enable_timing(self, on=True):
        self.lib.sgpu_enable_timing(self.ctx, 1 if on else 0)

    def last_ms(self, phase):
        return self.lib.sgpu_last_kernel_ms(self.ctx, phase.encode())
