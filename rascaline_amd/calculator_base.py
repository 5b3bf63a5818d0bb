"""Python side of the calculators, mirroring featomic's ``CalculatorBase``
(python/featomic/featomic/calculator_base.py:60-314) over the C API of
librascaline_amd.so."""

import ctypes
import json
from typing import List, Optional

import numpy as np

from . import _c_lib as C
from .systems import PAIR_DTYPE, _SystemAdapter, wrap_system
from .tensormap import Labels, TensorBlock, TensorMap


def _labels(raw):
    names = [raw.names[k].decode("utf8") for k in range(raw.size)]
    count = raw.count
    if count and raw.size:
        values = np.ctypeslib.as_array(raw.values, shape=(count * raw.size,)).reshape(count, raw.size).copy()
    else:
        values = np.zeros((count, raw.size), dtype=np.int32)
    return Labels(names, values)


def _radial_size(radial):
    if radial.get("type") == "Tabulated":
        return len(radial["points"][0]["values"]) if radial.get("points") else 0
    return int(radial.get("max_radial", 0)) + 1


def _basis_shape(basis):
    """(L, n) of the tensor-product shape a basis is computed in: the largest angular
    channel + 1 and the largest radial size."""
    if "by_angular" in basis:
        channels = basis["by_angular"]
        return max(int(k) for k in channels) + 1, max(_radial_size(r) for r in channels.values())
    return int(basis.get("max_angular", 0)) + 1, _radial_size(basis["radial"])


class _TensorMapOwner:
    """Keeps a library-owned mts_tensormap_t alive while numpy views of its arrays exist."""

    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        # at interpreter exit the library may already be unloaded
        if self.ptr and C._LIB is not None:
            C._LIB.mts_tensormap_free(self.ptr)
        self.ptr = None


def _read_block(ptr, owner):
    L = C.lib()
    arr = C.mts_array_t()
    C.check_mts(L.mts_block_data(ptr, ctypes.byref(arr)))
    shape_ptr = ctypes.POINTER(ctypes.c_size_t)()
    ndim = ctypes.c_size_t()
    C.check_mts(arr.shape(arr.ptr, ctypes.byref(shape_ptr), ctypes.byref(ndim)))
    shape = tuple(shape_ptr[k] for k in range(ndim.value))
    data_ptr = ctypes.POINTER(ctypes.c_double)()
    C.check_mts(arr.data(arr.ptr, ctypes.byref(data_ptr)))
    size = int(np.prod(shape)) if shape else 0
    if size:
        # a view of the library's array (no copy); the buffer keeps the TensorMap alive
        buf = (ctypes.c_double * size).from_address(ctypes.addressof(data_ptr.contents))
        buf._owner = owner
        values = np.frombuffer(buf, dtype=np.float64).reshape(shape)
    else:
        values = np.zeros(shape)
    labels = []
    for axis in range(len(shape)):
        raw = C.mts_labels_t()
        C.check_mts(L.mts_block_labels(ptr, axis, ctypes.byref(raw)))
        labels.append(_labels(raw))
        L.mts_labels_free(ctypes.byref(raw))
    block = TensorBlock(values, labels[0], labels[1:-1], labels[-1])
    names = ctypes.POINTER(ctypes.c_char_p)()
    count = ctypes.c_size_t()
    C.check_mts(L.mts_block_gradients_list(ptr, ctypes.byref(names), ctypes.byref(count)))
    for k in range(count.value):
        param = names[k]
        g = ctypes.c_void_p()
        C.check_mts(L.mts_block_gradient(ptr, param, ctypes.byref(g)))
        block.add_gradient(param.decode("utf8"), _read_block(g, owner))
    return block


def tensormap_from_ptr(ptr):
    """A Python TensorMap over a library-owned mts_tensormap_t: labels are copied, the
    block arrays are numpy views of the library's float64 arrays, which are released
    with the last of them."""
    L = C.lib()
    owner = _TensorMapOwner(ptr)
    raw = C.mts_labels_t()
    C.check_mts(L.mts_tensormap_keys(ptr, ctypes.byref(raw)))
    keys = _labels(raw)
    blocks = []
    for b in range(len(keys)):
        bp = ctypes.c_void_p()
        C.check_mts(L.mts_tensormap_block_by_id(ptr, ctypes.byref(bp), b))
        blocks.append(_read_block(bp, owner))
    return TensorMap(keys, blocks)


class _SelectionArgs:
    """ctypes views of the selections of one compute call, alive as long as this object."""

    def __init__(self):
        self._keep = []
        self._tensors = []

    def labels(self, labels):
        names = [n.encode("utf8") for n in labels.names]
        name_array = (ctypes.c_char_p * max(len(names), 1))(*names)
        values = np.ascontiguousarray(labels.values, dtype=np.int32)
        self._keep += [name_array, values]
        raw = C.mts_labels_t()
        raw.internal_ptr_ = None
        raw.names = ctypes.cast(name_array, ctypes.POINTER(ctypes.c_char_p))
        raw.values = values.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        raw.size = len(names)
        raw.count = values.shape[0]
        self._keep.append(raw)
        return raw

    def tensor(self, tensor):
        """A library-side TensorMap with the labels of `tensor` (zero data): the
        predefined selection (featomic's LabelsSelection::Predefined)."""
        L = C.lib()
        blocks = []
        for block in tensor.blocks():
            components = [self.labels(c) for c in block.components]
            comp_array = (C.mts_labels_t * max(len(components), 1))(*components)
            self._keep.append(comp_array)
            ptr = L.mts_block(C.mts_array_t(), self.labels(block.samples), comp_array, len(components),
                              self.labels(block.properties))
            if not ptr:
                for b in blocks:
                    L.mts_block_free(b)
                raise C.FeatomicError(L.mts_last_error().decode("utf8"), 1)
            blocks.append(ptr)
        array = (ctypes.c_void_p * max(len(blocks), 1))(*blocks)
        ptr = L.mts_tensormap(self.labels(tensor.keys), array, len(blocks))
        if not ptr:
            for b in blocks:
                L.mts_block_free(b)
            raise C.FeatomicError(L.mts_last_error().decode("utf8"), 1)
        self._tensors.append(ptr)
        return ptr

    def selection(self, value, out):
        if value is None:
            return
        if isinstance(value, Labels):
            out.subset = ctypes.pointer(self.labels(value))
        elif isinstance(value, TensorMap):
            out.predefined = self.tensor(value)
        else:
            raise TypeError(f"expected Labels or TensorMap as selection, got {type(value)}")

    def release(self):
        for ptr in self._tensors:
            C.lib().mts_tensormap_free(ptr)
        self._tensors = []


class CalculatorBase:
    """Base class of all calculators (featomic ``CalculatorBase``)."""

    def __init__(self, name: str, parameters: str):
        self._lib = C.lib()
        self._as_parameter_ = None
        ptr = self._lib.featomic_calculator(name.encode("utf8"), parameters.encode("utf8"))
        if not ptr:
            message = self._lib.featomic_last_error().decode("utf8")
            raise C.FeatomicError(message)
        self._as_parameter_ = ctypes.c_void_p(ptr)

    def __del__(self):
        if getattr(self, "_as_parameter_", None) is not None and C._LIB is not None:
            C._LIB.featomic_calculator_free(self._as_parameter_)
            self._as_parameter_ = None

    @property
    def name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        C.check(self._lib.featomic_calculator_name(self._as_parameter_, buf, 256))
        return buf.value.decode("utf8")

    @property
    def parameters(self) -> str:
        size = 1024
        while True:
            buf = ctypes.create_string_buffer(size)
            status = self._lib.featomic_calculator_parameters(self._as_parameter_, buf, size)
            if status == 254:
                size *= 2
                continue
            C.check(status)
            return buf.value.decode("utf8")

    @property
    def cutoffs(self) -> List[float]:
        ptr = ctypes.POINTER(ctypes.c_double)()
        count = ctypes.c_size_t()
        C.check(self._lib.featomic_calculator_cutoffs(self._as_parameter_, ctypes.byref(ptr), ctypes.byref(count)))
        return [ptr[k] for k in range(count.value)]

    def set_precision(self, bits: int):
        """Arithmetic of the GPU kernels: 64 (default) or 32 bits."""
        C.check(self._lib.featomic_amd_calculator_set_precision(self._as_parameter_, int(bits)))

    def set_fast_form(self, enabled: bool):
        """FP32 only: mixed-precision (bfloat16 hi + lo on MFMA) power-spectrum gradient
        contraction, within 1e-4 of the FP64 result."""
        C.check(self._lib.featomic_amd_calculator_set_fast_form(self._as_parameter_, 1 if enabled else 0))

    def train_radial_surrogate(self, **options) -> dict:
        """Train the learned radial integral on the GPU (SOAP calculators) and use it
        for later computations when the calculator's spline judges every (l, n)
        channel within ``tolerance``. Options: n_samples, steps, batch,
        learning_rate, ridge, tolerance, seed, n_validation (defaults from the
        library). Returns the report with per-(l, n) errors as [L][n] arrays."""
        opts = C.featomic_amd_surrogate_options_t()
        C.check(self._lib.featomic_amd_surrogate_options_default(ctypes.byref(opts)))
        for key, value in options.items():
            if not hasattr(opts, key):
                raise ValueError(f"unknown radial surrogate option '{key}'")
            setattr(opts, key, value)
        L, n = _basis_shape(json.loads(self.parameters)["basis"])
        value_errors = np.zeros((L, n))
        derivative_errors = np.zeros((L, n))
        report = C.featomic_amd_surrogate_report_t()
        report.value_errors = value_errors.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        report.derivative_errors = derivative_errors.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        report.errors_size = L * n
        C.check(self._lib.featomic_amd_calculator_train_radial_surrogate(self._as_parameter_, ctypes.byref(opts),
                                                                           ctypes.byref(report)))
        return {
            "accepted": bool(report.accepted),
            "adam_loss": report.adam_loss,
            "max_value_error": report.max_value_error,
            "max_derivative_error": report.max_derivative_error,
            "train_ms": report.train_ms,
            "value_errors": value_errors,
            "derivative_errors": derivative_errors,
        }

    def use_radial_surrogate(self, enabled: bool):
        """Use the accepted learned radial integral (True) or the spline / exact one."""
        C.check(self._lib.featomic_amd_calculator_use_radial_surrogate(self._as_parameter_, 1 if enabled else 0))

    def autotune(self, systems, gradients=("positions",), repeats=3, stream=None) -> dict:
        """Time the candidate launch shapes of the SOAP kernels on device-resident
        systems (a ``rascaline_amd.device.DeviceSystems``), keep the fastest for the
        following computations and return the timings of every candidate."""
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        names = [g.encode("utf8") for g in gradients]
        arr = (ctypes.c_char_p * max(len(names), 1))(*names)
        buf = ctypes.create_string_buffer(1 << 16)
        C.check(self._lib.featomic_amd_calculator_autotune(self._as_parameter_, ctypes.byref(systems.struct), arr,
                                                           len(names), ctypes.c_void_p(s.cuda_stream), int(repeats),
                                                           buf, len(buf)))
        return json.loads(buf.value.decode("utf8"))

    @property
    def tuning(self) -> dict:
        """Launch shapes in use: {"expansion_entries": 0, 2^k, -8 or -16, "gradient_window": -2, -1 or k}."""
        e, w = ctypes.c_int32(), ctypes.c_int32()
        C.check(self._lib.featomic_amd_calculator_get_tuning(self._as_parameter_, ctypes.byref(e), ctypes.byref(w)))
        return {"expansion_entries": e.value, "gradient_window": w.value}

    @tuning.setter
    def tuning(self, value: dict):
        C.check(self._lib.featomic_amd_calculator_set_tuning(self._as_parameter_, int(value.get("expansion_entries", 0)),
                                                             int(value.get("gradient_window", -1))))

    def radial_surrogate_parameters(self) -> np.ndarray:
        """Flat FP32 parameters of the trained network, then its output scales."""
        required = ctypes.c_int64()
        C.check(self._lib.featomic_amd_calculator_radial_surrogate_parameters(self._as_parameter_, None, 0,
                                                                                ctypes.byref(required)))
        out = np.zeros(required.value, dtype=np.float32)
        C.check(self._lib.featomic_amd_calculator_radial_surrogate_parameters(
            self._as_parameter_, out.ctypes.data_as(ctypes.c_void_p), out.size, ctypes.byref(required)))
        return out

    def compute(self, systems, *, gradients: Optional[List[str]] = None, use_native_system: bool = True,
                selected_samples=None, selected_properties=None, selected_keys=None) -> TensorMap:
        """Run the calculation (python/featomic/featomic/calculator_base.py:187-314)."""
        if not isinstance(systems, list):
            systems = [systems]
        adapters = [_SystemAdapter(wrap_system(s)) for s in systems]
        array = (C.featomic_system_t * len(adapters))(*[a.struct for a in adapters])
        gradients = list(gradients or [])
        grads = (ctypes.c_char_p * max(len(gradients), 1))(*[g.encode("utf8") for g in gradients])
        options = C.featomic_calculation_options_t()
        options.gradients = ctypes.cast(grads, ctypes.POINTER(ctypes.c_char_p))
        options.gradients_count = len(gradients)
        options.use_native_system = bool(use_native_system)
        # selections (calculator_base.py:230-285): Labels for a subset, TensorMap for a
        # predefined selection, Labels for the keys
        selections = _SelectionArgs()
        try:
            selections.selection(selected_samples, options.selected_samples)
            selections.selection(selected_properties, options.selected_properties)
            if selected_keys is not None:
                if not isinstance(selected_keys, Labels):
                    raise TypeError(f"expected Labels for selected_keys, got {type(selected_keys)}")
                options.selected_keys = ctypes.pointer(selections.labels(selected_keys))
            out = ctypes.c_void_p()
            C.check(self._lib.featomic_calculator_compute(self._as_parameter_, ctypes.byref(out), array, len(adapters),
                                                          options))
        finally:
            selections.release()
        if not out.value:
            return None
        return tensormap_from_ptr(out)


def neighbors(positions, cell, cutoff):
    """Half neighbour list computed on the GPU, as ``featomic_pair_t`` records."""
    L = C.lib()
    positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    ptr = ctypes.POINTER(C.featomic_pair_t)()
    count = ctypes.c_size_t()
    C.check(L.featomic_amd_neighbors(positions.ctypes.data, len(positions), cell.ctypes.data, float(cutoff),
                                     ctypes.byref(ptr), ctypes.byref(count)))
    try:
        n = count.value
        if n == 0:
            return np.zeros(0, dtype=PAIR_DTYPE)
        raw = ctypes.string_at(ptr, n * PAIR_DTYPE.itemsize)
        return np.frombuffer(raw, dtype=PAIR_DTYPE).copy()
    finally:
        L.featomic_amd_free(ptr)
