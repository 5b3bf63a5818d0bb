"""Device-resident inputs and outputs of ``featomic_amd_compute_device``.

`DeviceSystems` holds a batch of systems whose positions live in GPU memory;
`compute_device` runs a calculator on it and returns a `DeviceResult` whose
blocks are torch tensors viewing the calculator's device buffers (no copy).
The views stay valid until the next computation with the same calculator or its
release; clone them to keep them longer.

The layout is the one of the host TensorMap (featomic's output contract): one
block per key, values [samples][properties], position gradients
[rows][3][properties], sample labels (system, atom) and gradient labels
(sample within the block, system, atom).
"""

import ctypes

import numpy as np

from . import _c_lib as C


class DeviceSystems:
    """Systems given as (positions [n][3], types [n], cell [3][3]) tuples, with the
    positions copied once to device memory (float64, torch allocation). With
    `positions` (a contiguous float64 GPU tensor [total atoms, 3]), that tensor is
    used in place of the tuples' positions, without a copy; `cells` ([systems][3][3])
    replaces the tuples' cells."""

    def __init__(self, systems, device="cuda:0", positions=None, cells=None):
        import torch

        if positions is None:
            pos = np.concatenate([np.asarray(s[0], dtype=np.float64).reshape(-1, 3) for s in systems])
            self.positions = torch.from_numpy(pos).to(device)
        else:
            if positions.dtype != torch.float64 or not positions.is_cuda or not positions.is_contiguous():
                raise ValueError("positions must be a contiguous float64 GPU tensor")
            if positions.shape != (sum(len(s[1]) for s in systems), 3):
                raise ValueError("positions must hold [total atoms, 3] values")
            self.positions = positions
        self.n_atoms = int(self.positions.shape[0])
        self.types = np.ascontiguousarray(np.concatenate([np.asarray(s[1]) for s in systems]).astype(np.int32))
        if cells is None:
            cells = [s[2] for s in systems]
        self.cells = np.ascontiguousarray(np.stack([np.asarray(c, dtype=np.float64).reshape(9) for c in cells]))
        self.offsets = np.zeros(len(systems) + 1, dtype=np.int64)
        self.offsets[1:] = np.cumsum([len(s[1]) for s in systems])
        self.struct = C.featomic_amd_device_systems_t()
        self.struct.n_systems = len(systems)
        self.struct.atom_offsets = self.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        self.struct.positions = ctypes.c_void_p(self.positions.data_ptr())
        self.struct.types = self.types.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        self.struct.cells = self.cells.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class _DeviceArray:
    """A device pointer described through the CUDA array interface (version 3),
    which torch.as_tensor accepts without copying."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": tuple(int(s) for s in shape), "typestr": typestr,
                                         "data": (int(ptr), False), "version": 3, "strides": None}


def _view(ptr, shape, typestr, device):
    import torch

    if ptr is None or ptr == 0 or int(np.prod(shape)) == 0:
        dtype = {"<f4": torch.float32, "<f8": torch.float64, "<i4": torch.int32}[typestr]
        return torch.empty(shape, dtype=dtype, device=device)
    return torch.as_tensor(_DeviceArray(ptr, shape, typestr), device=device)


class DeviceResult:
    """Blocks of one device computation as torch views (see the module docstring)."""

    def __init__(self, out, device="cuda:0"):
        self.n_blocks = int(out.n_blocks)
        ks = int(out.key_size)
        self.keys = np.array([[out.keys[b * ks + k] for k in range(ks)] for b in range(self.n_blocks)], dtype=np.int32)
        self.n_samples = [int(out.n_samples[b]) for b in range(self.n_blocks)]
        self.n_gradient_rows = [int(out.n_gradient_rows[b]) for b in range(self.n_blocks)]
        self.properties = [int(out.block_properties[b]) for b in range(self.n_blocks)]
        self.precision = int(out.precision)
        self._typestr = "<f4" if self.precision == 32 else "<f8"
        self._item = 4 if self.precision == 32 else 8
        self._out = out
        self._device = device
        self._slab = [int(out.sample_label_offsets[b]) for b in range(self.n_blocks)] \
            if out.sample_label_offsets else None
        self._glab = [int(out.gradient_label_offsets[b]) for b in range(self.n_blocks)] \
            if out.gradient_label_offsets else None

    def value_offset(self, b):
        """Element offset of block b's values in the result's value buffer (the layout
        positions_vjp expects for dA/dX)."""
        return int(self._out.value_offsets[b])

    @property
    def value_size(self):
        """Elements of the value buffer: the end of the last block."""
        return max((self.value_offset(b) + self.n_samples[b] * self.properties[b] for b in range(self.n_blocks)),
                   default=0)

    def values(self, b):
        off = int(self._out.value_offsets[b])
        ptr = self._out.values + off * self._item if self._out.values else 0
        return _view(ptr, (self.n_samples[b], self.properties[b]), self._typestr, self._device)

    def gradients(self, b):
        if not self._out.gradients:
            raise ValueError("no position gradients were computed")
        off = int(self._out.gradient_offsets[b])
        return _view(self._out.gradients + off * self._item, (self.n_gradient_rows[b], 3, self.properties[b]),
                     self._typestr, self._device)

    def cell_gradients(self, b):
        """[samples][3 (xyz)][3 (abc)][properties] cell gradients of block b (the host
        TensorMap's "cell" gradient, one row per sample)."""
        if not self._out.cell_gradients:
            raise ValueError("no cell gradients were computed")
        off = int(self._out.cell_gradient_offsets[b])
        return _view(self._out.cell_gradients + off * self._item, (self.n_samples[b], 3, 3, self.properties[b]),
                     self._typestr, self._device)

    def sample_labels(self, b):
        """[samples][2] (system, atom) of block b, or None when the calculator does
        not produce labels."""
        if not self._out.sample_labels or self._slab is None:
            return None
        return _view(self._out.sample_labels + 8 * self._slab[b], (self.n_samples[b], 2), "<i4", self._device)

    def gradient_labels(self, b):
        """[rows][3] (sample within the block, system, atom) of block b, or None when
        the calculator does not produce labels."""
        if not self._out.gradient_labels or self._glab is None:
            return None
        return _view(self._out.gradient_labels + 12 * self._glab[b], (self.n_gradient_rows[b], 3), "<i4",
                     self._device)

    @property
    def total_rows(self):
        return sum(self.n_gradient_rows)


def compute_device(calculator, systems, gradients=(), stream=None):
    """Run `calculator` on the device-resident `systems` (a DeviceSystems); the
    work is queued on `stream` (a torch.cuda.Stream, or the current stream) and
    the result's tensors are ready once that stream has reached this point."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    names = [g.encode("utf8") for g in gradients]
    arr = (ctypes.c_char_p * max(len(names), 1))(*names)
    out = C.featomic_amd_device_output_t()
    C.check(C.lib().featomic_amd_compute_device(calculator._as_parameter_, ctypes.byref(systems.struct), arr,
                                                len(names), ctypes.c_void_p(s.cuda_stream), ctypes.byref(out)))
    return DeviceResult(out)


def positions_vjp(calculator, systems, grad_values, stream=None):
    """dA/dr [n_atoms, 3] = Σ_blocks Σ_rows (dX/dr)[row] · dA/dX[sample of row] for a SOAP
    calculator on device-resident `systems`, without forming the gradient rows.
    `grad_values` is a contiguous device tensor in the layout of the values of
    `compute_device(calculator, systems)` (all blocks, at their value offsets) and
    the calculator's precision."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    dtype = grad_values.dtype
    if not grad_values.is_cuda or not grad_values.is_contiguous():
        raise ValueError("grad_values must be a contiguous GPU tensor")
    out = torch.empty((systems.n_atoms, 3), dtype=dtype, device=grad_values.device)
    C.check(C.lib().featomic_amd_positions_vjp(calculator._as_parameter_, ctypes.byref(systems.struct),
                                               ctypes.c_void_p(grad_values.data_ptr()), ctypes.c_void_p(s.cuda_stream),
                                               ctypes.c_void_p(out.data_ptr())))
    return out


class BatchGraph:
    """One pass of SOAP calculators over many device batches captured into a single
    hipGraph (``featomic_amd_batch_graph_create``). Batch k runs on
    ``calculators[k % len(calculators)]``; the graph copies the batches' positions
    when it is created and every ``launch`` recomputes all batches from them. The
    calculators belong to the graph until ``close``.

        graph = BatchGraph([calc_a, calc_b], batches, ["positions"])
        graph.launch()                 # asynchronous, on torch's current stream
        graph.check()                  # waits; raises if a replay was inconsistent
        result = graph.output(k)       # DeviceResult of batch k (last batches of each lane)
        graph.close()
    """

    def __init__(self, calculators, batches, gradients=(), stream=None):
        import torch

        self._batches = list(batches)
        self._calculators = list(calculators)
        s = stream if stream is not None else torch.cuda.current_stream()
        calcs = (ctypes.c_void_p * len(self._calculators))(*[c._as_parameter_.value for c in self._calculators])
        structs = (C.featomic_amd_device_systems_t * max(len(self._batches), 1))()
        for k, b in enumerate(self._batches):
            structs[k] = b.struct
        names = [g.encode("utf8") for g in gradients]
        arr = (ctypes.c_char_p * max(len(names), 1))(*names)
        handle = ctypes.c_void_p()
        C.check(C.lib().featomic_amd_batch_graph_create(calcs, len(self._calculators), structs, len(self._batches), arr,
                                                        len(names), ctypes.c_void_p(s.cuda_stream),
                                                        ctypes.byref(handle)))
        self._handle = handle
        self._keep = (calcs, structs, arr)

    def launch(self, stream=None):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        C.check(C.lib().featomic_amd_batch_graph_launch(self._handle, ctypes.c_void_p(s.cuda_stream)))

    def check(self):
        C.check(C.lib().featomic_amd_batch_graph_check(self._handle))

    def output(self, batch):
        out = C.featomic_amd_device_output_t()
        C.check(C.lib().featomic_amd_batch_graph_output(self._handle, int(batch), ctypes.byref(out)))
        return DeviceResult(out)

    def close(self):
        if self._handle:
            C.check(C.lib().featomic_amd_batch_graph_free(self._handle))
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
