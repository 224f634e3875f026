"""ctypes binding of ``libneuwon_amd.so`` (declared in ``include/neuwon_amd.h``).

The library is loaded from the package directory only; a missing library raises ImportError-like
``RuntimeError`` at first use.  There is no CPU fallback for any entry point.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libneuwon_amd.so")

c_i32, c_i64, c_u32, c_f32, c_f64, c_vp = (ctypes.c_int32, ctypes.c_int64, ctypes.c_uint32,
                                           ctypes.c_float, ctypes.c_double, ctypes.c_void_p)


NW_PROPAGATOR_SYMMETRIC = 1
NW_PROPAGATOR_CHEBYSHEV = 2
NW_PROPAGATOR_RATIONAL = 4


class NwPropagatorStats(ctypes.Structure):
    _fields_ = [("num_cells", c_i64), ("num_rows", c_i64), ("nnz", c_i64), ("stored", c_i64),
                ("max_cell_rows", c_i32), ("max_row_nnz", c_i32), ("build_seconds", c_f64),
                ("device_bytes", c_i64), ("symmetric", c_i32), ("chebyshev", c_i32),
                ("mean_terms", c_f64), ("workgroup_threads", c_i32), ("row_groups", c_i32),
                ("rational", c_i32), ("max_levels", c_i32), ("mean_levels", c_f64)]


class NwHHParams(ctypes.Structure):
    _fields_ = [("dt_ms", c_f32), ("vmin", c_f32), ("vmax", c_f32), ("num_nodes", c_i32),
                ("e_na", c_f32), ("e_k", c_f32), ("e_l", c_f32), ("stim_mode", c_i32),
                ("stim_amp_a", c_f32), ("stim_seed", c_u32), ("poisson_threshold", c_u32),
                ("pulse_ticks", c_i32), ("ap_threshold", c_f32), ("conductances_valid", c_i32),
                ("step0", c_i64), ("block_threads", c_i32), ("gate_mode", c_i32), ("q10", c_f32),
                ("launch_mode", c_i32), ("cheb_launch", c_i32), ("reserved", c_i32)]


NW_LAUNCH_DIRECT, NW_LAUNCH_GRAPH, NW_LAUNCH_GRAPH_PREPARE = 0, 1, 2
NW_CHEB_LAUNCH_DEFAULT, NW_CHEB_LAUNCH_HALF, NW_CHEB_LAUNCH_STEP, NW_CHEB_LAUNCH_TWO_STEPS = 0, 1, 2, 3


class NwCaParams(ctypes.Structure):
    _fields_ = [("dt_ms", c_f32), ("kon", c_f32), ("koff", c_f32), ("btot", c_f32),
                ("cab_ref", c_f64), ("block_threads", c_i32), ("pad_", c_i32)]


NW_EVENT_SET, NW_EVENT_ADD, NW_EVENT_RECORD, NW_EVENT_TRACE = 0, 1, 2, 3
NW_DTYPE_F32, NW_DTYPE_U8, NW_DTYPE_I32 = 0, 1, 2


class NwEvent(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("dtype", c_i32), ("column", c_vp), ("row", c_i64),
                ("first_tick", c_i64), ("end_tick", c_i64), ("wave", c_vp), ("length", c_i64),
                ("first_index", c_i64), ("slot", c_i64), ("state", c_vp), ("alpha", c_f64),
                ("beta", c_f64), ("loop", c_i32), ("run_end", c_i32)]


NW_INPUT_LINEAR, NW_INPUT_LOG2 = 0, 1


class NwLtiInput(ctypes.Structure):
    _fields_ = [("values", c_vp), ("index", c_vp), ("mapping", c_i32), ("num_nodes", c_i32),
                ("minimum", c_f64), ("maximum", c_f64), ("scale", c_f64)]


# name -> (restype, argtypes)
SIGNATURES = {
    "nw_last_error": (ctypes.c_char_p, []),
    "nw_version": (ctypes.c_char_p, []),
    "nw_device_info": (c_i32, [c_i32, ctypes.c_char_p, c_i32, ctypes.POINTER(c_i32),
                               ctypes.POINTER(c_i32)]),
    "nw_propagator_build": (c_i32, [c_i32, c_i64, c_vp, c_vp, c_vp, c_vp, c_f64, c_f64,
                                    ctypes.POINTER(c_vp)]),
    "nw_propagator_build_ex": (c_i32, [c_i32, c_i64, c_vp, c_vp, c_vp, c_vp, c_f64, c_f64, c_i32,
                                       ctypes.POINTER(c_vp)]),
    "nw_propagator_stats_get": (c_i32, [c_vp, ctypes.POINTER(NwPropagatorStats)]),
    "nw_propagator_cell_weights": (c_i32, [c_vp, c_i64, c_vp]),
    "nw_propagator_cell_chebyshev": (c_i32, [c_vp, c_i64, ctypes.POINTER(c_i32),
                                             ctypes.POINTER(c_f64), c_vp]),
    "nw_chebyshev_exp_coefficients": (c_i32, [c_f64, c_f64, c_vp, c_i32]),
    "nw_rational_coefficients": (c_i32, [c_vp, c_i32]),
    "nw_propagator_destroy": (c_i32, [c_vp]),
    "nw_propagator_cell_f64": (c_i32, [c_i32, c_i32, c_vp, c_vp, c_vp, c_f64, c_f64,
                                       ctypes.POINTER(c_i64), c_vp, c_vp, c_vp]),
    "nw_propagator_cell_export": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_vp]),
    "nw_electric_half_step": (c_i32, [c_vp, c_f32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "nw_accumulate": (c_i32, [c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "nw_ap_detect": (c_i32, [c_i64, c_vp, c_vp, c_f32, c_vp, c_vp, c_vp]),
    "nw_lti_advance": (c_i32, [c_i32, c_i32, c_f32, c_f32, c_vp, c_i64, c_vp, c_vp, c_vp, c_i64,
                               c_f32, c_i32, c_vp, c_vp, c_vp, c_vp]),
    "nw_lti_advance_nd": (c_i32, [c_i32, c_i32, c_vp, c_vp, c_i64, c_vp, c_i64, c_f32, c_i32,
                                  c_vp, c_vp, c_vp, c_vp]),
    "nw_lti_table_create": (c_i32, [c_i32, c_i32, c_i32, c_f32, c_f32, c_vp, ctypes.POINTER(c_vp)]),
    "nw_lti_table_destroy": (c_i32, [c_vp]),
    "nw_lti_expm_batched": (c_i32, [c_i64, c_i32, c_vp, c_f64, c_i32, c_vp, c_vp]),
    "nw_lti_table_advance": (c_i32, [c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_f32, c_i32, c_vp,
                                     c_vp, c_vp, c_i32, c_vp]),
    "nw_hh_step": (c_i32, [c_vp, ctypes.POINTER(NwHHParams), c_i32, c_vp, c_vp, c_vp, c_vp, c_vp,
                           c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "nw_hh_adjoint": (c_i32, [c_vp, ctypes.POINTER(NwHHParams), c_i32, c_vp, c_vp, c_vp, c_vp,
                              c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                              c_vp, c_vp, c_vp, c_vp]),
    "nw_stim_hash": (c_u32, [c_u32, c_u32, c_u32]),
    "nw_ca_step": (c_i32, [c_vp, ctypes.POINTER(NwCaParams), c_i32, c_vp, c_vp, c_vp, c_vp]),
    "nw_weighted_sum": (c_i32, [c_i64, c_vp, c_vp, c_vp, c_vp, c_f64, c_vp, c_i32, c_vp]),
    "nw_weighted_sum_total": (c_i32, [c_i64, c_vp, c_vp, c_vp, c_vp, c_f64, c_i32, c_vp, c_vp]),
    "nw_clock_tick": (c_i32, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp]),
    "nw_trace_all": (c_i32, [c_i64, c_vp, c_vp, c_vp, c_f64, c_f64, c_vp]),
    "nw_hh_mechanism": (c_i32, [c_i64, c_i32, c_f32, c_f32, c_vp, c_i32, c_f32, c_f32, c_vp, c_vp,
                                c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
}

_lib = None
_lock = threading.RLock()


def load():
    """Load the library (once).  Raises RuntimeError when it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m neuwon_amd.build`")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
        return _lib


def lock():
    """The library lock: calls that touch shared device state are serialised with it."""
    return _lock


def check(rc):
    if rc != 0:
        msg = load().nw_last_error().decode(errors="replace")
        raise RuntimeError(f"neuwon_amd error {rc}: {msg}")
    return rc


def ptr(t):
    """Device (or host) address of a torch tensor / numpy array as a ctypes void pointer.  The
    entry points take dense arrays, so a strided view (a transpose, a column slice) is refused
    rather than read in the wrong order."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        if not t.is_contiguous():
            raise ValueError(f"tensor of shape {tuple(t.shape)} and strides {t.stride()} is not "
                             "contiguous")
        return ctypes.c_void_p(t.data_ptr())
    if not t.flags["C_CONTIGUOUS"]:
        raise ValueError(f"array of shape {t.shape} is not C-contiguous")
    return ctypes.c_void_p(t.ctypes.data)
