"""ctypes binding of libsfft.so (the C ABI in include/sfft.h).

The library is built in-tree (gpu-sfft_amd/libsfft.so).  There is no fallback: if the .so is
missing or fails to load, importing this module raises, naming the file and the build step.
"""
import ctypes as C
import os

try:
    # torch ships its own HIP runtime (torch/lib/libamdhip64.so, soname libamdhip64.so.7).
    # Loaded first, it is the runtime libsfft.so binds to; if libsfft.so came first it would
    # bring in /opt/rocm's copy and torch would load a second runtime beside it, which then
    # sees no device.  So torch (when installed) is imported before the library is opened.
    import torch  # noqa: F401
except ImportError:
    pass

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.dirname(_PKG_DIR)
# SFFT_LIBRARY: another build of the same sources (an instrumented or experimental one)
LIB_PATH = os.environ.get("SFFT_LIBRARY") or os.path.join(LIB_DIR, "libsfft.so")
CSRC_DIR = os.path.join(LIB_DIR, "csrc")


class LibraryNotBuilt(ImportError):
    pass


class complex_t(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double)]


class Node(C.Structure):
    _fields_ = [("key", C.c_uint), ("value", complex_t)]


class Params(C.Structure):
    _fields_ = [("n", C.c_int), ("B", C.c_int), ("B_thresh", C.c_int), ("W_RLFC", C.c_int),
                ("RLFC_loops", C.c_int), ("loops_thresh", C.c_int), ("loops_loc", C.c_int),
                ("loops", C.c_int), ("w", C.c_int), ("fold_mode", C.c_int),
                ("hit_capacity", C.c_int), ("front_end", C.c_int)]


class Draws(C.Structure):
    # host int32 arrays, passed as addresses (sfft_draws: const int *a, *ai, *offsets)
    _fields_ = [("a", C.c_void_p), ("ai", C.c_void_p), ("offsets", C.c_void_p)]


SFFT_OK = 0
SFFT_ERR_INVALID = 1
SFFT_ERR_ALLOC = 2
SFFT_ERR_HIP = 3
SFFT_ERR_HIT_OVERFLOW = 4
SFFT_ERR_NO_FILTER = 5
SFFT_ERR_NO_DEVICE = 6

SFFT_FP64 = 0
SFFT_FP32 = 1
SFFT_FOLD_REFERENCE = 0
SFFT_FOLD_ALL_TAPS = 1
SFFT_FRONT_AUTO = 0
SFFT_FRONT_GATHER = 1
SFFT_FRONT_STREAM = 2
SFFT_FRONT_COLUMNS = 3

# every symbol include/sfft.h declares (functions and data)
EXPORTED_FUNCTIONS = [
    "outer_loop", "restr_loc_freq_coeff", "samples_indices", "perm_filter_cutoff",
    "reverse_hash", "estimate_values", "timesmod", "samples_indices_noalloc", "make_multiple_t",
    "cudafft_experiment",
    "run_experiment", "cuShift", "sfft_srandom", "sfft_random",
    "sfft_status_string", "sfft_params_validate", "sfft_plan_create", "sfft_plan_destroy",
    "sfft_plan_workspace_bytes", "sfft_plan_set_filter", "sfft_execute_async", "sfft_execute",
    "sfft_plan_set_profiling", "sfft_plan_set_graph", "sfft_plan_stage_times",
    "sfft_plan_interval_times",
    "sfft_plan_debug_buffer", "sfft_plan_comb_key_order", "sfft_plan_front_end",
    "sfft_value_apply", "sfft_value_adjoint",
    "sfft_batch_create", "sfft_batch_destroy", "sfft_batch_plan", "sfft_batch_set_filter",
    "sfft_batch_set_graph", "sfft_batch_execute_async", "sfft_batch_execute",
    "sfft_batch_counts_device", "sfft_batch_tune", "sfft_batch_set_shape",
    "sfft_fft_device", "sfft_twiddle", "sfft_fft_plan_passes", "sfft_version",
    "sfft_device_count", "sfft_filter_taps", "sfft_make_filter", "sfft_make_filter_ex",
    "sfft_filter_window", "sfft_filter_window_device",
    "sfft_dft_any_device", "sfft_convert_c128_to_c64", "sfft_convert_c64_to_c128",
    "sfft_signal_spectrum", "sfft_signal_synthesize", "sfft_truth_spectrum", "sfft_metric",
]
EXPORTED_DATA = ["ALGORITHM1", "VERBOSE", "TIMING"]


def _load():
    if not os.path.exists(LIB_PATH):
        raise LibraryNotBuilt(
            f"{LIB_PATH} is missing: build it with `make -C {CSRC_DIR}` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as e:
        raise LibraryNotBuilt(f"cannot load {LIB_PATH}: {e}") from e
    _declare(lib)
    return lib


_p = C.c_void_p
_i = C.c_int
_d = C.c_double


def _declare(L):
    L.sfft_status_string.argtypes = [_i]
    L.sfft_status_string.restype = C.c_char_p
    L.sfft_params_validate.argtypes = [C.POINTER(Params)]
    L.sfft_params_validate.restype = _i
    L.sfft_plan_create.argtypes = [C.POINTER(Params), _i, _p, C.POINTER(_p)]
    L.sfft_plan_create.restype = _i
    L.sfft_plan_destroy.argtypes = [_p]
    L.sfft_plan_destroy.restype = None
    L.sfft_plan_workspace_bytes.argtypes = [_p]
    L.sfft_plan_workspace_bytes.restype = C.c_size_t
    L.sfft_plan_set_filter.argtypes = [_p, _p, _p]
    L.sfft_plan_set_filter.restype = _i
    L.sfft_execute_async.argtypes = [_p, _p, C.POINTER(Draws), _p, _p]
    L.sfft_execute_async.restype = _i
    L.sfft_execute.argtypes = [_p, _p, C.POINTER(Draws), _p, C.POINTER(_i)]
    L.sfft_execute.restype = _i
    L.sfft_plan_set_profiling.argtypes = [_p, _i]
    L.sfft_plan_set_profiling.restype = _i
    L.sfft_plan_set_graph.argtypes = [_p, _i]
    L.sfft_plan_set_graph.restype = _i
    L.sfft_plan_stage_times.argtypes = [_p, _p]
    L.sfft_plan_stage_times.restype = _i
    L.sfft_plan_interval_times.argtypes = [_p, _p]
    L.sfft_plan_interval_times.restype = _i
    L.sfft_plan_debug_buffer.argtypes = [_p, _i]
    L.sfft_plan_debug_buffer.restype = _p
    L.sfft_plan_comb_key_order.argtypes = [_p, C.POINTER(_i), C.POINTER(_i)]
    L.sfft_plan_comb_key_order.restype = _i
    L.sfft_plan_front_end.argtypes = [_p]
    L.sfft_plan_front_end.restype = _i
    L.sfft_value_apply.argtypes = [_p, _p, C.POINTER(Draws), _i, _p, _p, _p]
    L.sfft_value_apply.restype = _i
    L.sfft_value_adjoint.argtypes = [_p, _p, C.POINTER(Draws), _i, _p, _p, _p]
    L.sfft_value_adjoint.restype = _i
    L.sfft_batch_create.argtypes = [C.POINTER(Params), _i, _i, _p, C.POINTER(_p)]
    L.sfft_batch_create.restype = _i
    L.sfft_batch_destroy.argtypes = [_p]
    L.sfft_batch_destroy.restype = None
    L.sfft_batch_plan.argtypes = [_p, _i]
    L.sfft_batch_plan.restype = _p
    L.sfft_batch_set_filter.argtypes = [_p, _i, _p, _p]
    L.sfft_batch_set_filter.restype = _i
    L.sfft_batch_set_graph.argtypes = [_p, _i]
    L.sfft_batch_set_graph.restype = _i
    L.sfft_batch_execute_async.argtypes = [_p, _p, C.POINTER(Draws)]
    L.sfft_batch_execute_async.restype = _i
    L.sfft_batch_execute.argtypes = [_p, _p, C.POINTER(Draws), _p, _p]
    L.sfft_batch_execute.restype = _i
    L.sfft_batch_counts_device.argtypes = [_p]
    L.sfft_batch_counts_device.restype = _p
    L.sfft_batch_tune.argtypes = [_p, _p, C.POINTER(Draws), _i, _p]
    L.sfft_batch_tune.restype = _i
    L.sfft_batch_set_shape.argtypes = [_p, _i, _i, _i, _i, _i]
    L.sfft_batch_set_shape.restype = _i
    L.sfft_fft_device.argtypes = [_p, _p, _i, _i, _i, _i, _p]
    L.sfft_fft_device.restype = _i
    L.sfft_twiddle.argtypes = [C.c_uint64, C.c_uint64, _i, C.POINTER(_d), C.POINTER(_d)]
    L.sfft_twiddle.restype = None
    L.sfft_fft_plan_passes.argtypes = [_i, _p]
    L.sfft_fft_plan_passes.restype = _i
    L.sfft_version.restype = C.c_char_p
    L.sfft_device_count.restype = _i
    L.sfft_filter_taps.argtypes = [_d, _d, C.POINTER(_d)]
    L.sfft_filter_taps.restype = _i
    L.sfft_make_filter.argtypes = [_d, _d, _i, _i, _i, _p, _p, _p]
    L.sfft_make_filter.restype = _i
    L.sfft_make_filter_ex.argtypes = [_d, _d, _i, _i, _i, _i, _p, _p, _p, _p]
    L.sfft_make_filter_ex.restype = _i
    L.sfft_filter_window.argtypes = [_p, _i, _i, _p]
    L.sfft_filter_window.restype = None
    L.sfft_filter_window_device.argtypes = [_p, _i, _i, _p, _p]
    L.sfft_filter_window_device.restype = _i
    L.sfft_dft_any_device.argtypes = [_p, _i, _i, _p]
    L.sfft_dft_any_device.restype = _i
    L.sfft_convert_c128_to_c64.argtypes = [_p, _p, C.c_size_t, _p]
    L.sfft_convert_c128_to_c64.restype = _i
    L.sfft_convert_c64_to_c128.argtypes = [_p, _p, C.c_size_t, _p]
    L.sfft_convert_c64_to_c128.restype = _i
    L.sfft_signal_spectrum.argtypes = [_i, _i, C.c_uint64, _p, _p]
    L.sfft_signal_spectrum.restype = _i
    L.sfft_signal_synthesize.argtypes = [_i, _i, _p, _p, _d, C.c_uint64, _i, _p, _p]
    L.sfft_signal_synthesize.restype = _i
    L.sfft_truth_spectrum.argtypes = [_p, _i, _i, _p, _p]
    L.sfft_truth_spectrum.restype = _i
    L.sfft_metric.argtypes = [_p, _i, _i, _p, _p, _i, C.POINTER(_i), C.POINTER(_i),
                              C.POINTER(_d), C.POINTER(_d), C.POINTER(_d), _p]
    L.sfft_metric.restype = _i
    # reference entry points
    L.sfft_srandom.argtypes = [C.c_uint]
    L.sfft_srandom.restype = None
    L.sfft_random.argtypes = []
    L.sfft_random.restype = C.c_long
    L.timesmod.argtypes = [_i, _i, _i]
    L.timesmod.restype = _i
    L.outer_loop.argtypes = [_p, _i, _p, _p, _i, _i, _i, _i, _i, _i, _i, _i, C.POINTER(_i),
                             C.c_float]
    L.outer_loop.restype = C.POINTER(Node)
    L.restr_loc_freq_coeff.argtypes = [_p, _i, _i, _i, _p, C.POINTER(_d)]
    L.restr_loc_freq_coeff.restype = _i
    L.samples_indices.argtypes = [_p, _i, _i, _p, C.POINTER(_d)]
    L.samples_indices.restype = _i
    L.samples_indices_noalloc.argtypes = [_p, _i, _i, _p, C.POINTER(_d)]
    L.samples_indices_noalloc.restype = _i
    L.perm_filter_cutoff.argtypes = [_p, _i, _p, _i, _i, _i, _i, _i, _p, _p, C.POINTER(_d),
                                     C.POINTER(_d)]
    L.perm_filter_cutoff.restype = _i
    L.reverse_hash.argtypes = [_p, _i, _i, _i, _i, _i, _i, _p, _p, C.POINTER(_i),
                               C.POINTER(_d), _p, _i, _i, _i]
    L.reverse_hash.restype = _i
    L.estimate_values.argtypes = [_p, C.POINTER(_i), _p, _i, _i, _p, _i, _p, _i]
    L.estimate_values.restype = C.POINTER(Node)


lib = _load()
