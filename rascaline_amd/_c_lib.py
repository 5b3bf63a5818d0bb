"""Loading of librascaline_amd.so and its C prototypes.

The library is the product: there is no CPU fallback. If it is missing, build it
with ``python -m rascaline_amd.build``.
"""

import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(HERE, "librascaline_amd.so")

featomic_status_t = ctypes.c_int32


class mts_labels_t(ctypes.Structure):
    _fields_ = [
        ("internal_ptr_", ctypes.c_void_p),
        ("names", ctypes.POINTER(ctypes.c_char_p)),
        ("values", ctypes.POINTER(ctypes.c_int32)),
        ("size", ctypes.c_size_t),
        ("count", ctypes.c_size_t),
    ]


_ORIGIN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64))
_DATA = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
_SHAPE = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_size_t)),
                          ctypes.POINTER(ctypes.c_size_t))


class mts_array_t(ctypes.Structure):
    _fields_ = [
        ("ptr", ctypes.c_void_p),
        ("origin", _ORIGIN),
        ("data", _DATA),
        ("shape", _SHAPE),
        ("reshape", ctypes.c_void_p),
        ("swap_axes", ctypes.c_void_p),
        ("create", ctypes.c_void_p),
        ("copy", ctypes.c_void_p),
        ("destroy", ctypes.c_void_p),
        ("move_samples_from", ctypes.c_void_p),
    ]


class featomic_pair_t(ctypes.Structure):
    _fields_ = [
        ("first", ctypes.c_size_t),
        ("second", ctypes.c_size_t),
        ("distance", ctypes.c_double),
        ("vector", ctypes.c_double * 3),
        ("cell_shift_indices", ctypes.c_int32 * 3),
    ]


_SIZE_CB = ctypes.CFUNCTYPE(featomic_status_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t))
_TYPES_CB = ctypes.CFUNCTYPE(featomic_status_t, ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)))
_POSITIONS_CB = ctypes.CFUNCTYPE(featomic_status_t, ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
_CELL_CB = ctypes.CFUNCTYPE(featomic_status_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double))
_NEIGHBORS_CB = ctypes.CFUNCTYPE(featomic_status_t, ctypes.c_void_p, ctypes.c_double)
_PAIRS_CB = ctypes.CFUNCTYPE(featomic_status_t, ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(featomic_pair_t)),
                             ctypes.POINTER(ctypes.c_size_t))
_PAIRS_CONTAINING_CB = ctypes.CFUNCTYPE(featomic_status_t, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.POINTER(featomic_pair_t)), ctypes.POINTER(ctypes.c_size_t))


class featomic_system_t(ctypes.Structure):
    _fields_ = [
        ("user_data", ctypes.c_void_p),
        ("size", _SIZE_CB),
        ("types", _TYPES_CB),
        ("positions", _POSITIONS_CB),
        ("cell", _CELL_CB),
        ("compute_neighbors", _NEIGHBORS_CB),
        ("pairs", _PAIRS_CB),
        ("pairs_containing", _PAIRS_CONTAINING_CB),
    ]


class featomic_labels_selection_t(ctypes.Structure):
    _fields_ = [("subset", ctypes.POINTER(mts_labels_t)), ("predefined", ctypes.c_void_p)]


class featomic_calculation_options_t(ctypes.Structure):
    _fields_ = [
        ("gradients", ctypes.POINTER(ctypes.c_char_p)),
        ("gradients_count", ctypes.c_size_t),
        ("use_native_system", ctypes.c_bool),
        ("selected_samples", featomic_labels_selection_t),
        ("selected_properties", featomic_labels_selection_t),
        ("selected_keys", ctypes.POINTER(mts_labels_t)),
    ]


class featomic_amd_device_systems_t(ctypes.Structure):
    _fields_ = [
        ("n_systems", ctypes.c_size_t),
        ("atom_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("positions", ctypes.c_void_p),
        ("types", ctypes.POINTER(ctypes.c_int32)),
        ("cells", ctypes.POINTER(ctypes.c_double)),
    ]


class featomic_amd_device_output_t(ctypes.Structure):
    _fields_ = [
        ("n_blocks", ctypes.c_size_t),
        ("keys", ctypes.POINTER(ctypes.c_int32)),
        ("key_size", ctypes.c_size_t),
        ("n_samples", ctypes.POINTER(ctypes.c_int64)),
        ("n_gradient_rows", ctypes.POINTER(ctypes.c_int64)),
        ("value_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("gradient_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("block_properties", ctypes.POINTER(ctypes.c_int64)),
        ("values", ctypes.c_void_p),
        ("gradients", ctypes.c_void_p),
        ("sample_labels", ctypes.c_void_p),
        ("gradient_labels", ctypes.c_void_p),
        ("precision", ctypes.c_int32),
        ("sample_label_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("gradient_label_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("cell_gradients", ctypes.c_void_p),
        ("cell_gradient_offsets", ctypes.POINTER(ctypes.c_int64)),
    ]


class featomic_amd_surrogate_options_t(ctypes.Structure):
    _fields_ = [
        ("n_samples", ctypes.c_int64),
        ("steps", ctypes.c_int64),
        ("batch", ctypes.c_int32),
        ("learning_rate", ctypes.c_double),
        ("ridge", ctypes.c_double),
        ("tolerance", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("n_validation", ctypes.c_int64),
    ]


class featomic_amd_surrogate_report_t(ctypes.Structure):
    _fields_ = [
        ("accepted", ctypes.c_int32),
        ("adam_loss", ctypes.c_double),
        ("max_value_error", ctypes.c_double),
        ("max_derivative_error", ctypes.c_double),
        ("train_ms", ctypes.c_double),
        ("value_errors", ctypes.POINTER(ctypes.c_double)),
        ("derivative_errors", ctypes.POINTER(ctypes.c_double)),
        ("errors_size", ctypes.c_int64),
    ]


# every function the C headers declare: (name, restype, argtypes)
PROTOTYPES = [
    ("featomic_last_error", ctypes.c_char_p, []),
    ("featomic_set_logging_callback", featomic_status_t, [ctypes.c_void_p]),
    ("featomic_calculator", ctypes.c_void_p, [ctypes.c_char_p, ctypes.c_char_p]),
    ("featomic_calculator_free", featomic_status_t, [ctypes.c_void_p]),
    ("featomic_calculator_name", featomic_status_t, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    ("featomic_calculator_parameters", featomic_status_t, [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]),
    ("featomic_calculator_cutoffs", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), ctypes.POINTER(ctypes.c_size_t)]),
    ("featomic_calculator_compute", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(featomic_system_t), ctypes.c_size_t,
      featomic_calculation_options_t]),
    ("featomic_profiling_clear", featomic_status_t, []),
    ("featomic_profiling_enable", featomic_status_t, [ctypes.c_bool]),
    ("featomic_profiling_get", featomic_status_t, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
    ("mts_last_error", ctypes.c_char_p, []),
    ("mts_labels_free", ctypes.c_int32, [ctypes.POINTER(mts_labels_t)]),
    ("mts_tensormap_free", ctypes.c_int32, [ctypes.c_void_p]),
    ("mts_tensormap_keys", ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(mts_labels_t)]),
    ("mts_tensormap_block_by_id", ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    ("mts_block_labels", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(mts_labels_t)]),
    ("mts_block_data", ctypes.c_int32, [ctypes.c_void_p, ctypes.POINTER(mts_array_t)]),
    ("mts_block_gradient", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("mts_block", ctypes.c_void_p,
     [mts_array_t, mts_labels_t, ctypes.POINTER(mts_labels_t), ctypes.c_size_t, mts_labels_t]),
    ("mts_block_free", ctypes.c_int32, [ctypes.c_void_p]),
    ("mts_tensormap", ctypes.c_void_p, [mts_labels_t, ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    ("mts_block_gradients_list", ctypes.c_int32,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_char_p)), ctypes.POINTER(ctypes.c_size_t)]),
    ("featomic_amd_calculator_set_precision", featomic_status_t, [ctypes.c_void_p, ctypes.c_int32]),
    ("featomic_amd_calculator_set_fast_form", featomic_status_t, [ctypes.c_void_p, ctypes.c_int32]),
    ("featomic_amd_neighbors", featomic_status_t,
     [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.POINTER(featomic_pair_t)),
      ctypes.POINTER(ctypes.c_size_t)]),
    ("featomic_amd_free", None, [ctypes.c_void_p]),
    ("featomic_amd_calculator_autotune", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(featomic_amd_device_systems_t), ctypes.POINTER(ctypes.c_char_p), ctypes.c_size_t,
      ctypes.c_void_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_size_t]),
    ("featomic_amd_calculator_set_tuning", featomic_status_t, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    ("featomic_amd_calculator_get_tuning", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    ("featomic_amd_positions_vjp", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(featomic_amd_device_systems_t), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("featomic_amd_compute_device", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(featomic_amd_device_systems_t), ctypes.POINTER(ctypes.c_char_p), ctypes.c_size_t,
      ctypes.c_void_p, ctypes.POINTER(featomic_amd_device_output_t)]),
    ("featomic_amd_batch_graph_create", featomic_status_t,
     [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.POINTER(featomic_amd_device_systems_t), ctypes.c_size_t,
      ctypes.POINTER(ctypes.c_char_p), ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("featomic_amd_batch_graph_launch", featomic_status_t, [ctypes.c_void_p, ctypes.c_void_p]),
    ("featomic_amd_batch_graph_check", featomic_status_t, [ctypes.c_void_p]),
    ("featomic_amd_batch_graph_output", featomic_status_t,
     [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(featomic_amd_device_output_t)]),
    ("featomic_amd_batch_graph_free", featomic_status_t, [ctypes.c_void_p]),
    ("featomic_amd_last_timings", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    ("featomic_amd_surrogate_options_default", featomic_status_t, [ctypes.POINTER(featomic_amd_surrogate_options_t)]),
    ("featomic_amd_calculator_train_radial_surrogate", featomic_status_t,
     [ctypes.c_void_p, ctypes.POINTER(featomic_amd_surrogate_options_t), ctypes.POINTER(featomic_amd_surrogate_report_t)]),
    ("featomic_amd_calculator_use_radial_surrogate", featomic_status_t, [ctypes.c_void_p, ctypes.c_int32]),
    ("featomic_amd_calculator_radial_surrogate_parameters", featomic_status_t,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    ("featomic_amd_mlp_loss_gradient", featomic_status_t,
     [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
      ctypes.c_void_p]),
    ("featomic_amd_mlp_loss_gradient_noalloc", featomic_status_t,
     [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p]),
    ("featomic_amd_mlp_forward", featomic_status_t,
     [ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
]

_LIB = None


class FeatomicError(Exception):
    def __init__(self, message, status=None):
        super().__init__(message)
        self.status = status


def lib():
    """The loaded library; raises if it was not built (there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build it with `python -m rascaline_amd.build` (HIP, gfx950)"
            )
        # torch ships its own HIP runtime, which finds no GPU when this library's
        # runtime has initialised the device first: an imported torch sets up its
        # GPU context before the library is loaded
        torch = sys.modules.get("torch")
        if torch is not None and torch.cuda.is_available():
            torch.cuda.init()
        L = ctypes.CDLL(LIBRARY_PATH)
        for name, restype, argtypes in PROTOTYPES:
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _LIB = L
    return _LIB


def check(status):
    if status != 0:
        message = lib().featomic_last_error().decode("utf8", errors="replace")
        raise FeatomicError(message, status)


def check_mts(status):
    if status != 0:
        raise FeatomicError(lib().mts_last_error().decode("utf8", errors="replace"), status)
