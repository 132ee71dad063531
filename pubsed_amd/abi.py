"""ctypes mirror of include/sedona_gpu.h (the C-ABI of the transport hot path).

The structs here are plain descriptors of pointers and sizes; the caller
fills them with device pointers (GPU library) or host pointers (CPU oracle).
Field order and types must match the header exactly; tests/test_abi.py
checks the sizes against a C compiler.
"""
import ctypes as C

i32, i64, u32, u64, f64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
P = C.c_void_p

SGPU_OK = 0
SGPU_NOT_BRACKETED = 1
SGPU_MAX_ITER = 2
SGPU_ERR_ARG = -1
SGPU_ERR_HIP = -2
SGPU_ERR_SIZE = -3
SGPU_ERR_NOT_CONVERGED = -4

PHOTON, GAMMARAY = 0, 1
MOVING, STOPPED, ESCAPED, ABSORBED = 0, 1, 2, 3
LOC_FLEX, LOC_LIN, LOC_LOG, LOC_NONE = 0, 1, 2, 3
GRID_1D_SPHERE, GRID_2D_CYLN = 1, 2
PREC_FP64, PREC_MIXED, PREC_FP64_BATCHED = 0, 1, 2
NLTE_FP64, NLTE_FP32_REFINE = 0, 1


class Locate(C.Structure):
    _fields_ = [("n", i32), ("type", i32), ("min", f64), ("del_", f64), ("x", P)]


class Grid(C.Structure):
    _fields_ = [("kind", i32), ("n_zones", i32), ("t_now", f64), ("r_core", f64),
                ("r_out", Locate), ("x_out", Locate), ("z_out", Locate), ("vol", P)]


class ZoneState(C.Structure):
    _fields_ = [("n_zones", i32), ("n_elems", i32), ("rho", P), ("T_gas", P), ("X", P),
                ("e_gas", P), ("grey_opacity", P), ("elem_Z", P), ("elem_A", P)]


class Atoms(C.Structure):
    _fields_ = [("n_atoms", i32), ("max_levels", i32), ("n_levels", i32), ("n_ions", i32),
                ("lev_off", P), ("ion_off", P),
                ("lev_E", P), ("lev_g", P), ("lev_ion", P), ("ion_chi", P),
                ("atom_zero_eps", P), ("n_lines", i64), ("line_nu", P), ("line_gf", P),
                ("line_El", P), ("line_ion", P), ("line_bin", P), ("line_bin_off", P),
                ("n_blines", i64), ("bline_l", P), ("bline_u", P), ("bline_nu", P), ("bline_f", P),
                ("bline_A", P), ("bline_Bul", P), ("bline_Blu", P), ("bline_bin", P),
                ("bline_bin_off", P)]


class OpacityParams(C.Structure):
    _fields_ = [("t_now", f64), ("dt", f64), ("epsilon", f64), ("temp_min", f64),
                ("temp_max", f64), ("maximum_opacity", f64), ("line_velocity_width", f64),
                ("fleck_alpha", f64), ("use_electron_scattering", i32),
                ("use_fuzz_expansion", i32), ("omit_scattering", i32),
                ("omit_composition_decay", i32), ("radiative_eq", i32), ("first_step", i32),
                ("use_line_expansion", i32), ("use_free_free", i32), ("use_bound_free", i32),
                ("no_ground_recomb", i32), ("use_bound_bound", i32),
                ("minimum_extinction", f64)]


class OpacityOut(C.Structure):
    _fields_ = [("abs", P), ("scat", P), ("emis_cdf", P), ("L_thermal", P),
                ("planck_mean", P), ("rosseland_mean", P), ("n_elec", P),
                ("compton_opac", P), ("photoion_opac", P), ("eps_imc", P),
                ("zone_status", P), ("ion_frac", P), ("ion_part", P), ("lev_n", P)]


class Packets(C.Structure):
    _fields_ = [("capacity", i64), ("x", P), ("y", P), ("z", P), ("Dx", P), ("Dy", P),
                ("Dz", P), ("t", P), ("e", P), ("nu", P), ("ind", P), ("type", P),
                ("fate", P), ("draw", P), ("id", P)]


class CoreSource(C.Structure):
    _fields_ = [("r_core", f64), ("T_core", f64), ("L_core", f64), ("core_frequency", f64)]


class Tables(C.Structure):
    _fields_ = [("abs", P), ("scat", P), ("emis_cdf", P), ("eps_imc", P),
                ("compton_opac", P), ("photoion_opac", P)]


class Spectrum(C.Structure):
    _fields_ = [("time_grid", Locate), ("nu_grid", Locate), ("mu_grid", Locate),
                ("phi_grid", Locate), ("flux", P), ("click", P)]


class Tallies(C.Structure):
    _fields_ = [("e_abs", P), ("J_nu", P), ("fx_rad", P), ("fy_rad", P), ("fz_rad", P),
                ("fr_rad", P), ("L_radio_dep", P), ("L_radio_emit", P), ("e_rad", P),
                ("optical", Spectrum), ("gamma", Spectrum)]


class TransportParams(C.Structure):
    _fields_ = [("t_now", f64), ("dt", f64), ("seed", u64), ("steady_state", i32),
                ("radiative_eq", i32), ("boundary_in_reflect", i32),
                ("boundary_out_reflect", i32), ("store_Jnu", i32), ("omit_scattering", i32),
                ("precision", i32), ("max_events", i32), ("use_ddmc", i32),
                ("ddmc_tau_threshold", f64)]


class Counters(C.Structure):
    _fields_ = [("n_packets", u64), ("n_events", u64), ("n_binshift", u64),
                ("n_boundary", u64), ("n_scatter", u64), ("n_redist", u64),
                ("n_escaped", u64), ("n_absorbed", u64), ("n_stopped", u64),
                ("n_moving", u64)]


COUNTER_FIELDS = [f for f, _ in Counters._fields_]

PY_STRUCTS = {
    "sgpu_locate": Locate, "sgpu_grid": Grid, "sgpu_zone_state": ZoneState,
    "sgpu_atoms": Atoms, "sgpu_opacity_params": OpacityParams, "sgpu_opacity_out": OpacityOut,
    "sgpu_packets": Packets, "sgpu_tables": Tables, "sgpu_spectrum": Spectrum,
    "sgpu_tallies": Tallies, "sgpu_transport_params": TransportParams,
    "sgpu_counters": Counters, "sgpu_core_source": CoreSource,
}

# every extern "C" entry point declared in include/sedona_gpu.h
EXPORTED_SYMBOLS = [
    "sgpu_create", "sgpu_destroy", "sgpu_set_stream", "sgpu_get_stream", "sgpu_synchronize",
    "sgpu_last_error", "sgpu_alloc", "sgpu_free", "sgpu_memcpy_h2d", "sgpu_memcpy_d2h",
    "sgpu_last_kernel_ms", "sgpu_enable_timing", "sgpu_opacity_build",
    "sgpu_emit_radioactive", "sgpu_initialize_particles", "sgpu_emit_inner_source", "sgpu_propagate",
    "sgpu_compact",
    "sgpu_compact_noalloc", "sgpu_use_own_stream", "sgpu_graph_begin", "sgpu_graph_end", "sgpu_graph_launch",
    "sgpu_graph_nodes", "sgpu_emit_radioactive_dev", "sgpu_propagate_dev", "sgpu_compact_dev",
    "sgpu_atomic_rate",
    "sgpu_wipe_radiation",
    "sgpu_normalize_tallies", "sgpu_nlte_solve", "sgpu_nlte_batched", "sgpu_nlte_autotune", "sgpu_nlte_set_tuning",
    "sgpu_homologous_step", "sgpu_adjoint_forward", "sgpu_adjoint_backward",
    "sgpu_nlte_set_rates", "sgpu_nlte_populations", "sgpu_solve_eq_temperature",
    "sgpu_nlte_radiative_rates", "sgpu_opacity_build_nlte", "sgpu_alias_tables",
]


def ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def locate(n, typ, mn, dl, x):
    return Locate(n, typ, mn, dl, ptr(x))
