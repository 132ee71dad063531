"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU oracle is the checker of the parity tests and the CPU baseline of
bench.py; product code never imports this module.  OracleBackend mirrors the
method names of pubsed_amd.runtime.GpuBackend and works on a Problem whose
Space is the host.
"""
import ctypes
import os
import subprocess

import numpy as np

from pubsed_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        B = ctypes.byref
        P = ctypes.POINTER
        V, I, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        lib.or_uniform.restype = D
        lib.or_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
        lib.or_uniform_f32.restype = ctypes.c_float
        lib.or_uniform_f32.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]
        lib.or_philox4x32.argtypes = [V, V, V]
        lib.or_log_init.restype = I
        lib.or_log_init.argtypes = [D, D, D, V, I, P(D)]
        lib.or_locate.restype = I
        lib.or_locate.argtypes = [P(abi.Locate), D]
        lib.or_locate_within_bounds.restype = I
        lib.or_locate_within_bounds.argtypes = [P(abi.Locate), D]
        lib.or_opacity_build_nlte.restype = I
        lib.or_opacity_build_nlte.argtypes = [P(abi.Locate), P(abi.ZoneState), P(abi.Atoms),
                                              P(abi.OpacityParams), ctypes.c_void_p, ctypes.c_void_p,
                                              I, P(abi.OpacityOut)]
        lib.or_opacity_build.restype = I
        lib.or_opacity_build.argtypes = [P(abi.Locate), P(abi.ZoneState), P(abi.Atoms),
                                         P(abi.OpacityParams), P(abi.OpacityOut)]
        lib.or_solve_lte.argtypes = [P(abi.Atoms), I, D, D, V, V, V]
        lib.or_solve_state.restype = I
        lib.or_solve_state.argtypes = [P(abi.Atoms), I, V, V, D, D, D, P(D), V, V, V, P(I)]
        lib.or_emit_radioactive.restype = I
        lib.or_emit_radioactive.argtypes = [P(abi.Grid), P(abi.Locate), P(abi.ZoneState),
                                            P(abi.Tables), P(abi.TransportParams), I64, I,
                                            ctypes.c_uint64, P(abi.Packets), I64, P(abi.Tallies),
                                            P(I64)]
        lib.or_initialize_particles.restype = I
        lib.or_initialize_particles.argtypes = [P(abi.Grid), P(abi.Locate), P(abi.ZoneState), V, V,
                                                P(abi.TransportParams), I64, ctypes.c_uint64,
                                                P(abi.Packets), I64, P(I64)]
        lib.or_emit_inner_source.restype = I
        lib.or_emit_inner_source.argtypes = [P(abi.Grid), P(abi.Locate), P(abi.CoreSource),
                                             P(abi.TransportParams), I64, ctypes.c_uint64, P(abi.Packets),
                                             I64, P(I64)]
        lib.or_propagate.restype = I
        lib.or_propagate.argtypes = [P(abi.Grid), P(abi.Locate), P(abi.Tables),
                                     P(abi.TransportParams), P(abi.Packets), I64, P(abi.Tallies),
                                     P(abi.Counters), I]
        lib.or_normalize_tallies.restype = I
        lib.or_normalize_tallies.argtypes = [P(abi.Grid), P(abi.Locate), P(abi.TransportParams),
                                             P(abi.Tallies)]
        lib.or_compact.restype = I64
        lib.or_compact.argtypes = [P(abi.Packets), I64, P(I64)]
        lib.or_homologous_step.argtypes = [P(abi.Grid), V, V, V, V, V, D]
        lib.or_nlte_solve.restype = I
        lib.or_nlte_solve.argtypes = [I, V, V, V, V]
        lib.or_nlte_set_rates.restype = I
        lib.or_nlte_set_rates.argtypes = [P(abi.Atoms), I, D, D, V, V, V, V, I, I, V, V]
        lib.or_nlte_populations.argtypes = [P(abi.Atoms), I, V, V, V, V]
        lib.or_solve_lte.argtypes = [P(abi.Atoms), I, D, D, V, V, V]
        lib.or_voigt_profile.restype = D
        lib.or_voigt_profile.argtypes = [D, D]
        lib.or_photo_cs.restype = D
        lib.or_photo_cs.argtypes = [D, D, D, I, D]
        lib.or_nlte_radiative_rates.argtypes = [P(abi.Atoms), I, P(abi.Locate), V, D, I, D, V, V, V]
        lib.or_solve_eq_temperature.argtypes = [P(abi.Locate), I, V, V, V, D, D, I, V, V]
        lib.or_lu_decomp.restype = I
        lib.or_lu_decomp.argtypes = [I, V, V]
        lib.or_lu_solve.argtypes = [I, V, V, V, V]
        lib.or_klein_nishina.restype = D
        lib.or_klein_nishina.argtypes = [D]
        lib.or_blackbody_nu.restype = D
        lib.or_blackbody_nu.argtypes = [D, D]
        lib.or_get_zone.restype = I
        lib.or_get_zone.argtypes = [P(abi.Grid), V]
        lib.or_get_next_zone.restype = I
        lib.or_get_next_zone.argtypes = [P(abi.Grid), V, V, I, D, P(D)]
        lib.or_decay.restype = D
        lib.or_decay.argtypes = [I, V, V, V, D, P(D), I]
        lib.or_decay_composition.argtypes = [I, V, V, V, D]
        lib.or_lorentz.argtypes = [P(abi.Grid), V, V, P(D), P(D), I]
        lib.or_compton_sample.argtypes = [D, I64, ctypes.c_uint64, V, V]
        lib.or_adjoint_forward.restype = I
        lib.or_adjoint_forward.argtypes = [P(abi.Grid), P(abi.Locate), P(abi.Locate), V, V, V,
                                           P(abi.TransportParams), P(abi.Packets), I64, V, V, V, V, V]
        lib.or_adjoint_backward.restype = I
        lib.or_adjoint_backward.argtypes = [P(abi.Grid), P(abi.Locate), P(abi.Locate), V, V, V,
                                            P(abi.TransportParams), P(abi.Packets), I64, V, V, V, V]
        _lib = lib
    return _lib


def ptr(a):
    return a.ctypes.data if a is not None else None


class OracleBackend:
    """CPU restatement of the reference loops, same methods as GpuBackend."""

    def __init__(self, n_threads=0):
        self.lib = load()
        self.n_threads = n_threads
        self.counters = abi.Counters()

    def opacity_build(self, P):
        return self.lib.or_opacity_build(ctypes.byref(P.nu), ctypes.byref(P.zs),
                                         ctypes.byref(P.atoms) if P.atoms else None,
                                         ctypes.byref(P.op_params), ctypes.byref(P.op_out))

    def opacity_build_nlte(self, P, J_nu, atom_nlte, use_collisions=1):
        """The opacity build with the atoms flagged in atom_nlte solved in NLTE
        from J_nu [nz][n_nu] (numpy) at every n_e evaluation."""
        import numpy as np
        J = np.ascontiguousarray(J_nu, dtype=np.float64)
        fl = np.ascontiguousarray(atom_nlte, dtype=np.int32)
        return self.lib.or_opacity_build_nlte(ctypes.byref(P.nu), ctypes.byref(P.zs), ctypes.byref(P.atoms),
                                              ctypes.byref(P.op_params), J.ctypes.data_as(ctypes.c_void_p),
                                              fl.ctypes.data_as(ctypes.c_void_p), use_collisions,
                                              ctypes.byref(P.op_out))

    def emit(self, P, n_emit, id_base=0, offset=0, force_rproc=0):
        nout = ctypes.c_int64(0)
        self.lib.or_emit_radioactive(ctypes.byref(P.grid), ctypes.byref(P.nu), ctypes.byref(P.zs),
                                     ctypes.byref(P.tables), ctypes.byref(P.tparams), n_emit,
                                     force_rproc, id_base, ctypes.byref(P.packets), offset,
                                     ctypes.byref(P.tallies), ctypes.byref(nout))
        return nout.value

    def emit_inner_source(self, P, core, n_emit, id_base=0, offset=0):
        nout = ctypes.c_int64(0)
        self.lib.or_emit_inner_source(ctypes.byref(P.grid), ctypes.byref(P.nu), ctypes.byref(core),
                                      ctypes.byref(P.tparams), n_emit, id_base, ctypes.byref(P.packets),
                                      offset, ctypes.byref(nout))
        return nout.value

    def initialize_particles(self, P, n_init, id_base=0, offset=0):
        """transport::initialize_particles from P's e_rad tally; rewrites
        P's emissivity CDF with the zones' blackbody distributions."""
        nout = ctypes.c_int64(0)
        self.lib.or_initialize_particles(ctypes.byref(P.grid), ctypes.byref(P.nu), ctypes.byref(P.zs),
                                         ptr(P.a["e_rad"]), ptr(P.a["emis_cdf"]), ctypes.byref(P.tparams),
                                         n_init, id_base, ctypes.byref(P.packets), offset,
                                         ctypes.byref(nout))
        return nout.value

    def propagate(self, P, n, counters=None):
        c = counters if counters is not None else self.counters
        return self.lib.or_propagate(ctypes.byref(P.grid), ctypes.byref(P.nu), ctypes.byref(P.tables),
                                     ctypes.byref(P.tparams), ctypes.byref(P.packets), n,
                                     ctypes.byref(P.tallies), ctypes.byref(c), self.n_threads)

    def compact(self, P, n, scratch_pk=None):
        ne = ctypes.c_int64(0)
        k = self.lib.or_compact(ctypes.byref(P.packets), n, ctypes.byref(ne))
        return k, ne.value

    def wipe(self, P, wipe_spectra=1):
        for k in ("e_abs", "fx_rad", "fy_rad", "fz_rad", "fr_rad", "L_radio_dep", "L_radio_emit",
                  "e_rad", "J_nu"):
            P.a[k][:] = 0
        if wipe_spectra:
            for k in ("opt_flux", "opt_click", "gam_flux", "gam_click"):
                P.a[k][:] = 0

    def normalize(self, P):
        return self.lib.or_normalize_tallies(ctypes.byref(P.grid), ctypes.byref(P.nu),
                                             ctypes.byref(P.tparams), ctypes.byref(P.tallies))

    def solve_eq_temperature(self, P, set_Tgas_to_Trad=0, n_eval=None):
        """transport::solve_eq_temperature: P's T array from its e_abs tally."""
        self.lib.or_solve_eq_temperature(ctypes.byref(P.nu), P.nz, ptr(P.a["abs"]), ptr(P.a["e_abs"]),
                                         ptr(P.a["e_rad"]), P.op_params.temp_min, P.op_params.temp_max,
                                         set_Tgas_to_Trad, ptr(P.a["T"]), ptr(n_eval))

    def homologous_step(self, P, dt):
        e = (P.grid.t_now + dt) / P.grid.t_now
        a = P.a
        self.lib.or_homologous_step(ctypes.byref(P.grid), ptr(a["rho"]), ptr(a.get("r_out")),
                                    ptr(a.get("x_out")), ptr(a.get("z_out")), ptr(a["vol"]), e)
        P.grid.t_now = P.grid.t_now + dt

    def nlte_solve(self, n, rates, lte, b, pop=None):
        return self.lib.or_nlte_solve(n, ptr(rates), ptr(lte), ptr(b), ptr(pop))


def counters_dict(c):
    return {f: int(getattr(c, f)) for f in abi.COUNTER_FIELDS}
