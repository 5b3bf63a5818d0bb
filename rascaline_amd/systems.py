"""Atomistic systems passed to the calculators (featomic ``IntoSystem``).

featomic accepts ASE / chemfiles / PySCF systems through ``wrap_system``
(python/featomic/featomic/systems/__init__.py); none of these packages is a
dependency here, so any object with ``types``, ``positions`` and ``cell``
attributes (or ASE's ``numbers`` / ``get_positions()`` / ``cell``) is accepted.
"""

import ctypes

import numpy as np

from . import _c_lib as C


class System:
    """A system: atomic types [n], positions [n, 3] (Angstrom), cell [3, 3] whose
    rows are the cell vectors (all zero for a non-periodic system). An optional
    half neighbour list (``featomic_pair_t`` records as a numpy structured array)
    is used when computing with ``use_native_system=False``."""

    def __init__(self, types, positions, cell=None, pairs=None):
        self.types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        if cell is None:
            cell = np.zeros((3, 3))
        self.cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)
        if len(self.types) != len(self.positions):
            raise ValueError("types and positions must have the same length")
        self.pairs = pairs

    def __len__(self):
        return len(self.types)


def wrap_system(system):
    if isinstance(system, System):
        return system
    if hasattr(system, "types") and hasattr(system, "positions"):
        return System(system.types, system.positions, getattr(system, "cell", None))
    if hasattr(system, "numbers") and hasattr(system, "get_positions"):  # ASE-like
        cell = np.array(system.cell[:]) if any(system.pbc) else np.zeros((3, 3))
        return System(system.numbers, system.get_positions(), cell)
    if isinstance(system, (tuple, list)) and len(system) in (2, 3):
        return System(*system)
    raise TypeError(f"can not convert {type(system)} to a system")


PAIR_DTYPE = np.dtype([("first", np.uint64), ("second", np.uint64), ("distance", np.float64),
                       ("vector", np.float64, 3), ("cell_shift_indices", np.int32, 3), ("_pad", np.int32)])


class _SystemAdapter:
    """Exposes a System through the featomic_system_t vtable (featomic.h:167-241)."""

    def __init__(self, system):
        self.system = system
        self._types = system.types
        self._positions = system.positions
        self._cell = system.cell.reshape(9).copy()
        self._pairs = None
        self.callbacks = [
            C._SIZE_CB(self._size),
            C._TYPES_CB(self._get_types),
            C._POSITIONS_CB(self._get_positions),
            C._CELL_CB(self._get_cell),
            C._NEIGHBORS_CB(self._compute_neighbors),
            C._PAIRS_CB(self._get_pairs),
            C._PAIRS_CONTAINING_CB(self._pairs_containing),
        ]
        self.struct = C.featomic_system_t(None, *self.callbacks)

    def _size(self, _, size):
        size[0] = len(self._types)
        return 0

    def _get_types(self, _, out):
        out[0] = self._types.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        return 0

    def _get_positions(self, _, out):
        out[0] = self._positions.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return 0

    def _get_cell(self, _, out):
        for k in range(9):
            out[k] = self._cell[k]
        return 0

    def _compute_neighbors(self, _, cutoff):
        pairs = self.system.pairs
        if pairs is None:
            return -1
        self._pairs = np.ascontiguousarray(pairs, dtype=PAIR_DTYPE)
        return 0

    def _get_pairs(self, _, out, count):
        if self._pairs is None:
            return -1
        count[0] = len(self._pairs)
        out[0] = ctypes.cast(self._pairs.ctypes.data, ctypes.POINTER(C.featomic_pair_t))
        return 0

    def _pairs_containing(self, _, atom, out, count):
        return -1
