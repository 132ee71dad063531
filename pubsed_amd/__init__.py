"""pubsed_amd -- MI355X-native implicit Monte Carlo transport hot path of
Sedona (packet propagation, opacity/emissivity build, NLTE solves).

The compute path is libsedona_gpu.so (hand-written HIP for gfx950, C-ABI in
include/sedona_gpu.h).  This package is the host-side binding: ctypes
descriptors (abi), seeded workloads (workloads), device plumbing (runtime)
and the transport / GasState driver (transport).  There is no CPU fallback:
if the library is missing, loading fails loudly.
"""
import ctypes
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsedona_gpu.so")

_lib = None


class SgpuError(RuntimeError):
    pass


def _declare(lib):
    V, I, I32, I64, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "sgpu_create": (V, [I]),
        "sgpu_destroy": (None, [V]),
        "sgpu_set_stream": (I, [V, V]),
        "sgpu_get_stream": (V, [V]),
        "sgpu_synchronize": (I, [V]),
        "sgpu_last_error": (I, [ctypes.c_char_p, ctypes.c_size_t]),
        "sgpu_alloc": (I, [V, P(V), ctypes.c_size_t]),
        "sgpu_free": (I, [V, V]),
        "sgpu_memcpy_h2d": (I, [V, V, V, ctypes.c_size_t]),
        "sgpu_memcpy_d2h": (I, [V, V, V, ctypes.c_size_t]),
        "sgpu_last_kernel_ms": (D, [V, ctypes.c_char_p]),
        "sgpu_enable_timing": (I, [V, I]),
        "sgpu_opacity_build": (I, [V, P(abi.Locate), P(abi.ZoneState), P(abi.Atoms),
                                   P(abi.OpacityParams), P(abi.OpacityOut)]),
        "sgpu_emit_radioactive": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.ZoneState),
                                      P(abi.Tables), P(abi.TransportParams), I64, I32,
                                      ctypes.c_uint64, P(abi.Packets), I64, P(abi.Tallies),
                                      P(I64)]),
        "sgpu_initialize_particles": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.ZoneState), V, V,
                                          P(abi.TransportParams), I64, ctypes.c_uint64, P(abi.Packets),
                                          I64, P(I64)]),
        "sgpu_emit_inner_source": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.CoreSource),
                                       P(abi.TransportParams), I64, ctypes.c_uint64, P(abi.Packets), I64,
                                       P(I64)]),
        "sgpu_propagate": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.Tables),
                               P(abi.TransportParams), P(abi.Packets), I64, P(abi.Tallies), V]),
        "sgpu_compact": (I, [V, P(abi.Packets), I64, P(abi.Packets), P(I64), P(I64)]),
        "sgpu_compact_noalloc": (I, [V, P(abi.Packets), I64, P(abi.Packets), P(I64), P(I64)]),
        "sgpu_use_own_stream": (I, [V]),
        "sgpu_graph_begin": (I, [V]),
        "sgpu_graph_end": (I, [V, I32]),
        "sgpu_graph_launch": (I, [V]),
        "sgpu_graph_nodes": (I64, [V]),
        "sgpu_emit_radioactive_dev": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.ZoneState), P(abi.Tables),
                                          P(abi.TransportParams), I64, I32, ctypes.c_uint64, P(abi.Packets), V,
                                          P(abi.Tallies)]),
        "sgpu_propagate_dev": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.Tables), P(abi.TransportParams),
                                   P(abi.Packets), V, P(abi.Tallies), V]),
        "sgpu_compact_dev": (I, [V, P(abi.Packets), V, P(abi.Packets), V]),
        "sgpu_wipe_radiation": (I, [V, I32, I32, P(abi.Tallies), I32]),
        "sgpu_normalize_tallies": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.TransportParams),
                                       P(abi.Tallies)]),
        "sgpu_nlte_solve": (I, [V, I32, V, V, V, V, V]),
        "sgpu_nlte_batched": (I, [V, I32, V, V, V, V, V, V, V, I32, D, V]),
        "sgpu_nlte_autotune": (I, [V, I32, V, V, V, V, V, V, V, V, ctypes.c_char_p,
                                   ctypes.c_size_t]),
        "sgpu_nlte_set_tuning": (I, [V, I32, I32, I32, I32]),
        "sgpu_atomic_rate": (I, [V, V, ctypes.c_int64, ctypes.c_int64, P(D)]),
        "sgpu_homologous_step": (I, [V, P(abi.Grid), V, V, V, V, V, D]),
        "sgpu_nlte_set_rates": (I, [V, P(abi.Atoms), I32, V, V, V, V, V, V, V, I32, I32, V, V, V,
                                    V, V]),
        "sgpu_nlte_populations": (I, [V, P(abi.Atoms), I32, V, V, V, V, V, V]),
        "sgpu_solve_eq_temperature": (I, [V, P(abi.Locate), I32, V, V, V, D, D, I32, V, V]),
        "sgpu_nlte_radiative_rates": (I, [V, P(abi.Atoms), P(abi.Locate), I32, V, V, V, V, D, V, V, V, V]),
        "sgpu_alias_tables": (I, [V, V, I32, I32, V]),
        "sgpu_opacity_build_nlte": (I, [V, P(abi.Locate), P(abi.ZoneState), P(abi.Atoms),
                                        P(abi.OpacityParams), V, V, I32, I32, P(abi.OpacityOut), V]),
        "sgpu_adjoint_forward": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.Locate), V, V, V,
                                     P(abi.TransportParams), P(abi.Packets), I64, V, V, V]),
        "sgpu_adjoint_backward": (I, [V, P(abi.Grid), P(abi.Locate), P(abi.Locate), V, V, V,
                                      P(abi.TransportParams), P(abi.Packets), I64, V, V, V, V]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load_library(path=None):
    """Load libsedona_gpu.so (built by pubsed_amd.build / __graft_entry__.build)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    # SGPU_LIBRARY: load another build of the library (A/B timing of builds)
    p = path or os.environ.get("SGPU_LIBRARY") or LIB_PATH
    if not os.path.exists(p):
        raise SgpuError("libsedona_gpu.so not built (%s): run python -m pubsed_amd.build" % p)
    lib = _declare(ctypes.CDLL(p))
    if path is None:
        _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load_library().sgpu_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what=""):
    """Raise SgpuError for negative status codes; return non-negative ones."""
    if rc < 0:
        raise SgpuError("%s failed (%d): %s" % (what, rc, last_error()))
    return rc
