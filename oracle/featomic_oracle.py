"""FP64 CPU oracle for the featomic SOAP / LODE calculators.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker. The product package
(``rascaline_amd``) never imports it.

The numerical work lives in the C library built from this directory
(``liboracle.so``, see ``oracle.h``); this module assembles the metatensor-style
metadata (keys, samples, gradient samples, properties) with numpy, restating
the reference's label builders:

* SOAP spherical expansion keys / samples / gradient samples:
  ``featomic/src/calculators/soap/spherical_expansion.rs:662-803``,
  ``featomic/src/labels/keys.rs:61-107``,
  ``featomic/src/labels/samples/atom_centered.rs:31-150``;
* gradient row values: ``soap/spherical_expansion.rs:426-604``;
* SOAP power spectrum: ``soap/power_spectrum.rs:82-767`` and ``labels/keys.rs:101-170``;
* LODE: ``lode/spherical_expansion.rs:463-880``, ``labels/samples/long_range.rs:25-78``;
* NeighborList: ``calculators/neighbor_list.rs:108-620``;
* SphericalExpansionByPair: ``soap/spherical_expansion_pair.rs:222-300, 420-840``.

Outputs are plain dictionaries::

    {"keys_names": [...], "keys": int32[K, k],
     "blocks": [{"values", "samples_names", "samples", "components": [(name, values)],
                 "properties_names", "properties",
                 "gradients": {parameter: {same fields}}}]}
"""

import ctypes
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class OrcPair(ctypes.Structure):
    _fields_ = [
        ("first", ctypes.c_size_t),
        ("second", ctypes.c_size_t),
        ("distance", ctypes.c_double),
        ("vector", ctypes.c_double * 3),
        ("shift", ctypes.c_int32 * 3),
    ]


class SoapHypers(ctypes.Structure):
    _fields_ = [
        ("cutoff", ctypes.c_double),
        ("smoothing", ctypes.c_int),
        ("smoothing_width", ctypes.c_double),
        ("gaussian_width", ctypes.c_double),
        ("has_scaling", ctypes.c_int),
        ("scaling_scale", ctypes.c_double),
        ("scaling_rate", ctypes.c_double),
        ("scaling_exponent", ctypes.c_double),
        ("center_atom_weight", ctypes.c_double),
        ("max_angular", ctypes.c_int),
        ("n_radial", ctypes.c_int),
        ("spline_accuracy", ctypes.c_double),
        ("channel_n", ctypes.POINTER(ctypes.c_int)),
        ("channel_points", ctypes.POINTER(ctypes.c_int)),
        ("channel_positions", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
        ("channel_values", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
        ("channel_derivatives", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
    ]


class LodeHypers(ctypes.Structure):
    _fields_ = [
        ("smearing", ctypes.c_double),
        ("exponent", ctypes.c_int),
        ("center_atom_weight", ctypes.c_double),
        ("max_angular", ctypes.c_int),
        ("n_radial", ctypes.c_int),
        ("radius", ctypes.c_double),
        ("k_cutoff", ctypes.c_double),
        ("spline_accuracy", ctypes.c_double),
    ]


_LIB = None


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, i32, sz, vp = ctypes.c_double, ctypes.c_int32, ctypes.c_size_t, ctypes.c_void_p
        L.orc_gamma.restype = d
        L.orc_gamma.argtypes = [d]
        L.orc_ln_gamma.restype = d
        L.orc_ln_gamma.argtypes = [d]
        L.orc_hyp1f1.restype = d
        L.orc_hyp1f1.argtypes = [d, d, d]
        L.orc_hyp2f1.restype = d
        L.orc_hyp2f1.argtypes = [d, d, d, d]
        L.orc_expi.restype = d
        L.orc_expi.argtypes = [d]
        L.orc_spherical_harmonics.argtypes = [ctypes.c_int, vp, vp, vp]
        L.orc_symmetric_eigen.argtypes = [ctypes.c_int, vp, vp]
        L.orc_neighbors.restype = ctypes.c_int
        L.orc_neighbors.argtypes = [vp, sz, vp, d, ctypes.POINTER(ctypes.POINTER(OrcPair)), ctypes.POINTER(sz)]
        L.orc_free.argtypes = [vp]
        L.orc_soap_radial_new.restype = vp
        L.orc_soap_radial_new.argtypes = [d, d, ctypes.c_int, ctypes.c_int, d]
        L.orc_lode_radial_new.restype = vp
        L.orc_lode_radial_new.argtypes = [d, ctypes.c_int, ctypes.c_int, d, d]
        L.orc_radial_compute.argtypes = [vp, ctypes.c_int, d, vp, vp]
        L.orc_radial_n_knots.restype = ctypes.c_int
        L.orc_radial_n_knots.argtypes = [vp, ctypes.c_int]
        L.orc_radial_knots.argtypes = [vp, ctypes.c_int, vp, vp, vp]
        L.orc_radial_orthonormalization.argtypes = [vp, vp]
        L.orc_radial_free.argtypes = [vp]
        L.orc_soap_new.restype = vp
        L.orc_soap_new.argtypes = [ctypes.POINTER(SoapHypers)]
        L.orc_soap_free.argtypes = [vp]
        L.orc_soap_self_contribution.argtypes = [vp, vp]
        L.orc_soap_pair_contribution.argtypes = [vp, d, vp, ctypes.c_int, vp, vp]
        L.orc_cutoff_smoothing.restype = d
        L.orc_cutoff_smoothing.argtypes = [vp, d]
        L.orc_cutoff_smoothing_gradient.restype = d
        L.orc_cutoff_smoothing_gradient.argtypes = [vp, d]
        L.orc_soap_se_system.restype = ctypes.c_int
        L.orc_soap_se_system.argtypes = [vp, sz, vp, sz, vp, ctypes.c_int, vp, vp, vp, vp, vp, vp]
        L.orc_soap_ps_batch.restype = ctypes.c_int64
        L.orc_soap_ps_batch.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, vp, ctypes.c_int, ctypes.POINTER(d)]
        L.orc_ps_gradient_rows.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, d, vp, vp, vp, vp, vp, vp]
        L.orc_lode_new.restype = vp
        L.orc_lode_new.argtypes = [ctypes.POINTER(LodeHypers)]
        L.orc_lode_free.argtypes = [vp]
        L.orc_k_vectors.restype = ctypes.c_int
        L.orc_k_vectors.argtypes = [vp, d, vp]
        L.orc_lode_density_fourier.restype = d
        L.orc_lode_density_fourier.argtypes = [vp, d]
        L.orc_lode_center_contribution.argtypes = [vp, vp]
        L.orc_lode_k0_contribution.restype = ctypes.c_int
        L.orc_lode_k0_contribution.argtypes = [vp, vp]
        L.orc_lode_system.restype = ctypes.c_int
        L.orc_lode_system_centers.restype = ctypes.c_int
        L.orc_lode_system_centers.argtypes = [vp, sz, vp, vp, vp, ctypes.c_int, vp, vp, sz, vp, vp, vp]
        L.orc_lode_system.argtypes = [vp, sz, vp, vp, vp, ctypes.c_int, vp, vp, vp, vp]
        L.orc_mlp_n_params.restype = ctypes.c_int64
        L.orc_mlp_n_params.argtypes = [ctypes.c_int]
        L.orc_mlp_forward.argtypes = [ctypes.c_int, ctypes.c_int64, vp, vp, vp]
        L.orc_mlp_loss_gradient.restype = d
        L.orc_mlp_loss_gradient.argtypes = [ctypes.c_int, ctypes.c_int64, vp, vp, vp, vp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None



# ----------------------------------------------------------------------------
# radial-integral surrogate network (orc_mlp.c)

MLP_HIDDEN = 64


def mlp_n_params(n_out):
    return int(lib().orc_mlp_n_params(n_out))


def mlp_forward(params, inputs, n_out):
    """FP64 forward of the 1 -> 64 -> 64 -> n_out SiLU network."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    inputs = np.ascontiguousarray(inputs, dtype=np.float64)
    assert params.size == mlp_n_params(n_out)
    out = np.zeros((inputs.size, n_out))
    lib().orc_mlp_forward(n_out, inputs.size, _ptr(params), _ptr(inputs), _ptr(out))
    return out


def mlp_loss_gradient(params, inputs, targets, n_out):
    """Mean squared error over samples and outputs, and its gradient with
    respect to the flat parameter vector (FP64)."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    inputs = np.ascontiguousarray(inputs, dtype=np.float64)
    targets = np.ascontiguousarray(targets, dtype=np.float64)
    assert params.size == mlp_n_params(n_out) and targets.shape == (inputs.size, n_out)
    grad = np.zeros(params.size)
    loss = lib().orc_mlp_loss_gradient(n_out, inputs.size, _ptr(params), _ptr(inputs), _ptr(targets), _ptr(grad))
    return loss, grad


def radial_table(radial, r):
    """Exact (or splined) radial integral at every r: values and derivatives,
    each [len(r)][L][n]."""
    L = radial.max_angular + 1
    v = np.zeros((len(r), L, radial.n))
    g = np.zeros((len(r), L, radial.n))
    for i, x in enumerate(r):
        for l in range(L):
            v[i, l], g[i, l] = radial.compute(l, x)
    return v, g

# ----------------------------------------------------------------------------
# systems and hyper-parameters


class System:
    """positions [n, 3] (f64), types [n] (i32), cell [3, 3] (rows = cell vectors, zero = no pbc)."""

    def __init__(self, positions, types, cell):
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self.types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        self.cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(3, 3)

    def __len__(self):
        return len(self.types)


def load_input(path):
    """Read a ``*-input.json`` fixture (featomic/tests/data/save_data.py:15-52)."""
    with open(path) as fd:
        data = json.load(fd)
    systems = [System(s["positions"], s["types"], np.array(s["cell"]).reshape(3, 3)) for s in data["systems"]]
    return systems, data["hyperparameters"]


def soap_hypers(h):
    """JSON hyper-parameters -> SoapHypers (soap/spherical_expansion_pair.rs:37-60)."""
    out = SoapHypers()
    out.cutoff = float(h["cutoff"]["radius"])
    sm = h["cutoff"]["smoothing"]
    if sm["type"] == "ShiftedCosine":
        out.smoothing = 1
        out.smoothing_width = float(sm["width"])
    elif sm["type"] == "Step":
        out.smoothing = 0
    else:
        raise ValueError(sm)
    dens = h["density"]
    assert dens["type"] in ("Gaussian", "DiracDelta")
    out.gaussian_width = float(dens["width"]) if dens["type"] == "Gaussian" else 0.0
    sc = dens.get("scaling")
    if sc is not None:
        out.has_scaling = 1
        out.scaling_scale = float(sc["scale"])
        out.scaling_rate = float(sc["rate"])
        out.scaling_exponent = float(sc["exponent"])
    out.center_atom_weight = float(dens.get("center_atom_weight", 1.0))
    basis = h["basis"]
    acc = basis.get("spline_accuracy", 1e-8)
    out.spline_accuracy = -1.0 if acc is None else float(acc)
    channels = basis_channels(basis)
    out.max_angular = max(l for l, _ in channels)
    out.n_radial = max(_radial_size(r) for _, r in channels)
    one_gto = basis["type"] == "TensorProduct" and basis["radial"]["type"] == "Gto"
    if not one_gto:
        # per-channel radial integrals (radial_integral/mod.rs:138-183)
        L = out.max_angular + 1
        by_l = dict(channels)
        n_arr = (ctypes.c_int * L)()
        p_arr = (ctypes.c_int * L)()
        pos_arr = (ctypes.POINTER(ctypes.c_double) * L)()
        val_arr = (ctypes.POINTER(ctypes.c_double) * L)()
        der_arr = (ctypes.POINTER(ctypes.c_double) * L)()
        keep = [n_arr, p_arr, pos_arr, val_arr, der_arr]
        for l in range(L):
            radial = by_l.get(l)
            if radial is None:
                continue
            n_arr[l] = _radial_size(radial)
            if radial["type"] == "Tabulated":
                # points sorted by position (tabulated.rs:206-208)
                pts = sorted(radial["points"], key=lambda p: p["position"])
                pos = np.ascontiguousarray([p["position"] for p in pts], dtype=np.float64)
                val = np.ascontiguousarray([p["values"] for p in pts], dtype=np.float64)
                der = np.ascontiguousarray([p["derivatives"] for p in pts], dtype=np.float64)
                keep += [pos, val, der]
                p_arr[l] = len(pts)
                pos_arr[l] = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                val_arr[l] = val.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                der_arr[l] = der.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        out.channel_n = ctypes.cast(n_arr, ctypes.POINTER(ctypes.c_int))
        out.channel_points = ctypes.cast(p_arr, ctypes.POINTER(ctypes.c_int))
        out.channel_positions = ctypes.cast(pos_arr, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
        out.channel_values = ctypes.cast(val_arr, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
        out.channel_derivatives = ctypes.cast(der_arr, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)))
        out._keep = keep
    return out


def _radial_size(radial):
    """SoapRadialBasis::size (basis/radial/mod.rs:40-48)."""
    if radial["type"] == "Gto":
        return int(radial["max_radial"]) + 1
    return len(radial["points"][0]["values"])


def basis_channels(basis):
    """[(o3_lambda, radial basis)] of a SphericalExpansionBasis, ascending
    (basis/mod.rs:26-40): 0..max_angular sharing one radial basis, or the Explicit
    basis' by_angular map."""
    if basis.get("type", "TensorProduct") == "TensorProduct" and "by_angular" not in basis:
        if "max_angular" not in basis:
            return [(0, basis["radial"])]  # the radial spectrum's basis
        return [(l, basis["radial"]) for l in range(int(basis["max_angular"]) + 1)]
    return sorted((int(k), v) for k, v in basis["by_angular"].items())


def lode_hypers(h):
    out = LodeHypers()
    dens = h["density"]
    assert dens["type"] == "SmearedPowerLaw"
    out.smearing = float(dens["smearing"])
    out.exponent = int(dens["exponent"])
    out.center_atom_weight = float(dens.get("center_atom_weight", 1.0))
    basis = h["basis"]
    out.max_angular = int(basis["max_angular"])
    out.n_radial = int(basis["radial"]["max_radial"]) + 1
    out.radius = float(basis["radial"]["radius"])
    k_cutoff = h.get("k_cutoff")
    out.k_cutoff = -1.0 if k_cutoff is None else float(k_cutoff)
    acc = basis.get("spline_accuracy", 1e-8)
    out.spline_accuracy = -1.0 if acc is None else float(acc)
    return out


# ----------------------------------------------------------------------------
# primitives


def neighbors(positions, cell, cutoff):
    """Half neighbour list, sorted by (first, second, shift). Returns a dict of arrays."""
    L = lib()
    positions = np.ascontiguousarray(positions, dtype=np.float64)
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    pp = ctypes.POINTER(OrcPair)()
    n = ctypes.c_size_t()
    L.orc_neighbors(_ptr(positions), len(positions), _ptr(cell), float(cutoff), ctypes.byref(pp), ctypes.byref(n))
    count = n.value
    dtype = np.dtype([("first", np.uint64), ("second", np.uint64), ("distance", np.float64),
                      ("vector", np.float64, 3), ("shift", np.int32, 3), ("pad", np.int32)])
    assert dtype.itemsize == ctypes.sizeof(OrcPair)
    if count:
        buf = ctypes.string_at(pp, count * dtype.itemsize)
        arr = np.frombuffer(buf, dtype=dtype).copy()
    else:
        arr = np.zeros(0, dtype=dtype)
    L.orc_free(ctypes.cast(pp, ctypes.c_void_p))
    return arr


def neighbor_list(systems, params, gradients=()):
    """NeighborList calculator (featomic/src/calculators/neighbor_list.rs): the pair
    vectors of the half list of each system, relabelled per (first, second) atom-type
    key. Half list: one sample per pair, stored from the lower type to the higher one
    (atoms swapped, shift and vector negated when the pair is inverted,
    neighbor_list.rs:202-365); full list: each pair in both directions
    (neighbor_list.rs:416-620); self pairs (i, i, 0, 0, 0) with a zero vector after
    each system's pairs. Position gradients: -I for the first atom, +I for the second,
    none for a true self pair (neighbor_list.rs:108-127); a pair between an atom and
    its own periodic image gets one zero row."""
    cutoff = float(params["cutoff"])
    full = bool(params["full_neighbor_list"])
    self_pairs = bool(params.get("self_pairs", False))
    per_system = [neighbors(s.positions, s.cell, cutoff) for s in systems]
    keys = set()
    for s, pairs in zip(systems, per_system):
        for p in pairs:
            a, b = int(s.types[p["first"]]), int(s.types[p["second"]])
            if full:
                keys.update({(a, b), (b, a)})
            else:
                keys.add((min(a, b), max(a, b)))
        if self_pairs:
            keys.update((int(t), int(t)) for t in s.types)
    keys = sorted(keys)
    samples = {k: [] for k in keys}
    vectors = {k: [] for k in keys}

    def add(key, system, first, second, shift, vector):
        samples[key].append([system, int(first), int(second)] + [int(x) for x in shift])
        vectors[key].append([float(x) for x in vector])

    for si, (s, pairs) in enumerate(zip(systems, per_system)):
        for p in pairs:
            i, j = int(p["first"]), int(p["second"])
            a, b = int(s.types[i]), int(s.types[j])
            shift, vec = p["shift"], p["vector"]
            if full:
                add((a, b), si, i, j, shift, vec)
                add((b, a), si, j, i, -shift, -vec)
            elif a <= b:
                add((a, b), si, i, j, shift, vec)
            else:
                add((b, a), si, j, i, -shift, -vec)
        if self_pairs:
            for i, t in enumerate(s.types):
                add((int(t), int(t)), si, i, i, (0, 0, 0), (0.0, 0.0, 0.0))

    names = ["system", "first_atom", "second_atom", "cell_shift_a", "cell_shift_b", "cell_shift_c"]
    pair_xyz = ("pair_xyz", np.arange(3, dtype=np.int32).reshape(3, 1))
    blocks = []
    for k in keys:
        smp = samples[k]
        values = np.asarray(vectors[k], dtype=np.float64).reshape(len(smp), 3, 1)
        blk = _block(values, names, smp, [pair_xyz], ["distance"], [[1]])
        if "positions" in gradients:
            rows, gvals = [], []
            for r, (system, first, second, sa, sb, sc) in enumerate(smp):
                if first == second and (sa, sb, sc) == (0, 0, 0):
                    continue
                if first == second:
                    rows.append([r, system, first])
                    gvals.append(np.zeros((3, 3)))
                    continue
                rows += [[r, system, first], [r, system, second]]
                gvals += [-np.eye(3), np.eye(3)]
            gv = np.asarray(gvals, dtype=np.float64).reshape(len(rows), 3, 3, 1)
            blk["gradients"]["positions"] = _block(gv, ["sample", "system", "atom"], rows, [_xyz(), pair_xyz],
                                                   ["distance"], [[1]])
        blocks.append(blk)
    return {"keys_names": ["first_atom_type", "second_atom_type"],
            "keys": np.asarray(keys, dtype=np.int32).reshape(len(keys), 2), "blocks": blocks}


def spherical_harmonics(lmax, direction, gradients=True):
    direction = np.ascontiguousarray(direction, dtype=np.float64)
    L2 = (lmax + 1) ** 2
    values = np.zeros(L2)
    grads = np.zeros((3, L2)) if gradients else None
    lib().orc_spherical_harmonics(lmax, _ptr(direction), _ptr(values), _ptr(grads))
    return values, grads


class Radial:
    def __init__(self, handle, max_angular, n):
        self.handle = handle
        self.max_angular = max_angular
        self.n = n

    @staticmethod
    def soap(cutoff, width, max_angular, n_radial, spline_accuracy=None):
        h = lib().orc_soap_radial_new(cutoff, width, max_angular, n_radial, -1.0 if spline_accuracy is None else spline_accuracy)
        if not h:
            raise ValueError("radial integral construction failed")
        return Radial(h, max_angular, n_radial)

    @staticmethod
    def lode(radius, max_angular, n_radial, k_cutoff, spline_accuracy=None):
        h = lib().orc_lode_radial_new(radius, max_angular, n_radial, k_cutoff, -1.0 if spline_accuracy is None else spline_accuracy)
        if not h:
            raise ValueError("radial integral construction failed")
        return Radial(h, max_angular, n_radial)

    def compute(self, l, r):
        v = np.zeros(self.n)
        g = np.zeros(self.n)
        lib().orc_radial_compute(self.handle, l, float(r), _ptr(v), _ptr(g))
        return v, g

    def knots(self, l):
        k = lib().orc_radial_n_knots(self.handle, l)
        pos = np.zeros(k)
        val = np.zeros((k, self.n))
        der = np.zeros((k, self.n))
        lib().orc_radial_knots(self.handle, l, _ptr(pos), _ptr(val), _ptr(der))
        return pos, val, der

    def orthonormalization(self):
        out = np.zeros((self.n, self.n))
        lib().orc_radial_orthonormalization(self.handle, _ptr(out))
        return out

    def __del__(self):
        if self.handle and _LIB is not None:
            _LIB.orc_radial_free(self.handle)
            self.handle = None


# ----------------------------------------------------------------------------
# metadata helpers


def _block(values, samples_names, samples, components, properties_names, properties):
    return {
        "values": values,
        "samples_names": list(samples_names),
        "samples": np.asarray(samples, dtype=np.int32).reshape(len(samples), len(samples_names)),
        "components": components,
        "properties_names": list(properties_names),
        "properties": np.asarray(properties, dtype=np.int32).reshape(len(properties), len(properties_names)),
        "gradients": {},
    }


def _species(systems):
    return np.array(sorted({int(t) for s in systems for t in s.types}), dtype=np.int32)


def _xyz():
    return ("xyz", np.arange(3, dtype=np.int32).reshape(3, 1))


class _SystemData:
    """Neighbour list + per-atom neighbour sets for one system."""

    def __init__(self, system, cutoff):
        self.system = system
        self.pairs = neighbors(system.positions, system.cell, cutoff)
        n = len(system)
        # pairs_containing(atom): the other atom of each pair (self image -> atom itself)
        self.neighbors = [[] for _ in range(n)]
        for p, pair in enumerate(self.pairs):
            f, s = int(pair["first"]), int(pair["second"])
            self.neighbors[f].append((s, p))
            if s != f:
                self.neighbors[s].append((f, p))
            else:
                self.neighbors[s].append((f, p))


# ----------------------------------------------------------------------------
# SOAP spherical expansion


class SoapComputation:
    """Runs the C oracle once per system and keeps the dense results."""

    def __init__(self, systems, hypers, gradients=()):
        self.hypers = hypers
        self.h = soap_hypers(hypers)
        self.lmax = self.h.max_angular
        self.N = self.h.n_radial
        # (l, n_l) of the basis' angular channels
        self.channels = [(l, _radial_size(r)) for l, r in basis_channels(hypers["basis"])]
        self.L2 = (self.lmax + 1) ** 2
        self.species = _species(systems)
        self.S = len(self.species)
        self.gradients = set(gradients)
        handle = lib().orc_soap_new(ctypes.byref(self.h))
        if not handle:
            raise ValueError("invalid SOAP hyper-parameters")
        self.data = []
        try:
            for system in systems:
                sd = _SystemData(system, self.h.cutoff)
                n, P = len(system), len(sd.pairs)
                shape = (n, self.S, self.L2, self.N)
                sd.values = np.zeros(shape)
                sd.pair_grads = np.zeros((max(P, 1), 3, self.L2, self.N)) if "positions" in self.gradients else None
                sd.self_grads = np.zeros((n, self.S, 3, self.L2, self.N)) if "positions" in self.gradients else None
                sd.strain = np.zeros((n, self.S, 3, 3, self.L2, self.N)) if "strain" in self.gradients else None
                sd.cell_g = np.zeros((n, self.S, 3, 3, self.L2, self.N)) if "cell" in self.gradients else None
                pairs = np.ascontiguousarray(sd.pairs)
                status = lib().orc_soap_se_system(
                    handle, n, _ptr(system.types), P, _ptr(pairs) if P else None, self.S, _ptr(self.species),
                    _ptr(sd.values), _ptr(sd.pair_grads), _ptr(sd.self_grads), _ptr(sd.strain), _ptr(sd.cell_g))
                assert status == 0
                self.data.append(sd)
        finally:
            lib().orc_soap_free(handle)

    def tindex(self, t):
        return int(np.searchsorted(self.species, t))

    def se_position_gradient(self, sd, center, atom, nt_index, l):
        """Gradient of c_center^{nt} w.r.t. r_atom, [3, 2l+1, N] (spherical_expansion.rs:476-521)."""
        sl = slice(l * l, (l + 1) * (l + 1))
        if center == atom:
            return sd.self_grads[center, nt_index, :, sl, :]
        out = np.zeros((3, 2 * l + 1, self.N))
        m1l = -1.0 if l % 2 else 1.0
        for (other, p) in sd.neighbors[center]:
            if other != atom:
                continue
            pair = sd.pairs[p]
            factor = 1.0 if int(pair["first"]) == center else -m1l
            out += factor * sd.pair_grads[p, :, sl, :]
        return out


    def se_gradient_table(self, sd):
        """Table of the spherical-expansion position-gradient rows of one system,
        [entries][3][L2][N]: entry 0 is zero, then the self gradients of every (atom,
        type) from `self_base`, then one entry per (center, atom != center) with the sum
        over pair images (spherical_expansion.rs:476-521), indexed by `pair_index`."""
        cached = getattr(sd, "gradient_table", None)
        if cached is not None:
            return cached
        n = len(sd.system)
        sign = np.concatenate([np.full(2 * l + 1, -1.0 if l % 2 else 1.0) for l in range(self.lmax + 1)])
        pair_index = {}
        order = []
        for p, pair in enumerate(sd.pairs):
            f, s = int(pair["first"]), int(pair["second"])
            if f == s:
                continue
            for key in ((f, s), (s, f)):
                if key not in pair_index:
                    pair_index[key] = 1 + n * self.S + len(order)
                    order.append(key)
        table = np.zeros((1 + n * self.S + len(order), 3, self.L2, self.N))
        table[1:1 + n * self.S] = sd.self_grads.reshape(n * self.S, 3, self.L2, self.N)
        for p, pair in enumerate(sd.pairs):
            f, s = int(pair["first"]), int(pair["second"])
            if f == s:
                continue
            g = sd.pair_grads[p]
            table[pair_index[(f, s)]] += g
            table[pair_index[(s, f)]] -= sign[None, :, None] * g
        sd.gradient_table = (table, 1, pair_index)
        return sd.gradient_table

    def se_position_gradient_all(self, sd, center, atom, nt_index):
        """se_position_gradient for every l at once, [3, L2, N]; the per-(center, atom)
        sums over pair images are tabulated once per system."""
        if center == atom:
            return sd.self_grads[center, nt_index]
        table = getattr(sd, "pair_table", None)
        if table is None:
            sign = np.concatenate([np.full(2 * l + 1, -1.0 if l % 2 else 1.0) for l in range(self.lmax + 1)])
            table = {}
            for p, pair in enumerate(sd.pairs):
                f, s = int(pair["first"]), int(pair["second"])
                if f == s:
                    continue
                g = sd.pair_grads[p]
                table[(f, s)] = table[(f, s)] + g if (f, s) in table else g.copy()
                inv = -sign[None, :, None] * g
                table[(s, f)] = table[(s, f)] + inv if (s, f) in table else inv
            sd.pair_table = table
        out = table.get((center, atom))
        return out if out is not None else np.zeros((3, self.L2, self.N))


def soap_spherical_expansion(systems, hypers, gradients=()):
    comp = SoapComputation(systems, hypers, gradients)
    lmax, N = comp.lmax, comp.N

    # keys: CenterSingleNeighborsTypesKeys (labels/keys.rs:69-106)
    pairs_types = set()
    for sd in comp.data:
        t = sd.system.types
        for pair in sd.pairs:
            a, b = int(t[int(pair["first"])]), int(t[int(pair["second"])])
            pairs_types.add((a, b))
            pairs_types.add((b, a))
        for x in t:
            pairs_types.add((int(x), int(x)))
    keys = []
    blocks = []
    for (ct, nt) in sorted(pairs_types):
        # samples: AtomCenteredSamples, Single filters (atom_centered.rs:73-98)
        samples = []
        for si, sd in enumerate(comp.data):
            t = sd.system.types
            for i in range(len(t)):
                if t[i] != ct:
                    continue
                ok = nt == ct
                for (other, _) in sd.neighbors[i]:
                    if t[other] == nt:
                        ok = True
                if ok:
                    samples.append((si, i))
        # gradient samples (atom_centered.rs:103-150)
        grad_samples = []
        for s_i, (si, i) in enumerate(samples):
            sd = comp.data[si]
            t = sd.system.types
            nbrs = set()
            if t[i] == nt:
                nbrs.add(i)
            for (other, _) in sd.neighbors[i]:
                if t[other] == nt:
                    nbrs.add(other)
                    nbrs.add(i)
            for j in sorted(nbrs):
                grad_samples.append((s_i, si, j))
        nti = comp.tindex(nt)
        for l, nl in comp.channels:
            keys.append((l, 1, ct, nt))
            sl = slice(l * l, (l + 1) * (l + 1))
            values = np.zeros((len(samples), 2 * l + 1, nl))
            for s_i, (si, i) in enumerate(samples):
                values[s_i] = comp.data[si].values[i, nti, sl, :nl]
            mu = ("o3_mu", np.arange(-l, l + 1, dtype=np.int32).reshape(-1, 1))
            ns = np.arange(nl).reshape(-1, 1)
            blk = _block(values, ["system", "atom"], samples, [mu], ["n"], ns)
            if "positions" in comp.gradients:
                gv = np.zeros((len(grad_samples), 3, 2 * l + 1, nl))
                for r, (s_i, si, j) in enumerate(grad_samples):
                    i = samples[s_i][1]
                    gv[r] = comp.se_position_gradient(comp.data[si], i, j, nti, l)[..., :nl]
                blk["gradients"]["positions"] = _block(gv, ["sample", "system", "atom"], grad_samples, [_xyz(), mu], ["n"], ns)
            for param, attr, c0 in (("cell", "cell_g", "abc"), ("strain", "strain", "xyz_1")):
                if param not in comp.gradients:
                    continue
                gv = np.zeros((len(samples), 3, 3, 2 * l + 1, nl))
                for s_i, (si, i) in enumerate(samples):
                    gv[s_i] = getattr(comp.data[si], attr)[i, nti, :, :, sl, :nl]
                if param == "cell":
                    comps = [_xyz(), ("abc", np.arange(3).reshape(3, 1)), mu]
                else:
                    comps = [("xyz_2", np.arange(3).reshape(3, 1)), ("xyz_1", np.arange(3).reshape(3, 1)), mu]
                blk["gradients"][param] = _block(gv, ["sample"], np.arange(len(samples)).reshape(-1, 1), comps, ["n"], ns)
            blocks.append(blk)
    return {"keys_names": ["o3_lambda", "o3_sigma", "center_type", "neighbor_type"],
            "keys": np.array(keys, dtype=np.int32).reshape(len(keys), 4), "blocks": blocks}


# ----------------------------------------------------------------------------
# SOAP power spectrum


def soap_spherical_expansion_by_pair(systems, hypers, gradients=()):
    """SphericalExpansionByPair (featomic/src/calculators/soap/spherical_expansion_pair.rs):
    one sample per directed pair, the contribution of the second atom to the density
    around the first.

    * keys (o3_lambda, 1, first_atom_type, second_atom_type): the full-neighbour-list
      type pairs with self pairs (neighbor_list.rs:388-414), each with every l
      (spherical_expansion_pair.rs:560-583), in that order;
    * samples: the full neighbour list's samples (neighbor_list.rs:416-489), with the
      self pairs (i, i, 0, 0, 0) only in the l = 0 blocks (:585-637);
    * values: f Y_lm I_nl of the pair vector (orc_soap_pair_contribution); the
      reversed pair gets (-1)^l times the values and -(-1)^l times the vector
      gradients (PairContribution::inverse_pair, :127-155); self pairs hold the self
      contribution at l = 0 (:222-300);
    * position gradients: -g for the first atom, +g for the second, one row (their
      sum) for a pair of an atom with its own image, none for a self pair (:639-663,
      :450-493); strain [xyz_2, xyz_1] with data v[a] g[b]; cell [xyz, abc] with data
      shift[a] g[b] (:495-540).
    """
    h = soap_hypers(hypers)
    lmax, N = h.max_angular, h.n_radial
    L2 = (lmax + 1) ** 2
    cutoff = h.cutoff
    do_grad = bool(gradients)
    handle = lib().orc_soap_new(ctypes.byref(h))
    if not handle:
        raise ValueError("invalid SOAP hyper-parameters")
    try:
        self_c = np.zeros(N)
        lib().orc_soap_self_contribution(handle, _ptr(self_c))
        per_system = [neighbors(s.positions, s.cell, cutoff) for s in systems]
        type_pairs = set()
        for s, pairs in zip(systems, per_system):
            for p in pairs:
                a, b = int(s.types[p["first"]]), int(s.types[p["second"]])
                type_pairs.update({(a, b), (b, a)})
            type_pairs.update((int(t), int(t)) for t in s.types)
        type_pairs = sorted(type_pairs)
        # per type pair: [labels, values [L2][N], vector gradients [3][L2][N] or None, vector]
        samples = {tp: [] for tp in type_pairs}
        m1l = np.array([(-1.0) ** l for l in range(lmax + 1) for _ in range(2 * l + 1)])[:, None]
        for si, (s, pairs) in enumerate(zip(systems, per_system)):
            for p in pairs:
                i, j = int(p["first"]), int(p["second"])
                a, b = int(s.types[i]), int(s.types[j])
                vec = np.ascontiguousarray(p["vector"], dtype=np.float64)
                shift = [int(x) for x in p["shift"]]
                v = np.zeros((L2, N))
                g = np.zeros((3, L2, N))
                lib().orc_soap_pair_contribution(handle, float(p["distance"]), _ptr(vec), int(do_grad), _ptr(v),
                                                 _ptr(g) if do_grad else None)
                samples[(a, b)].append(([si, i, j] + shift, v, g, vec.copy()))
                samples[(b, a)].append(([si, j, i] + [-x for x in shift], m1l * v, -m1l[None] * g, -vec))
            for i, t in enumerate(s.types):
                v = np.zeros((L2, N))
                v[0] = self_c
                samples[(int(t), int(t))].append(([si, i, i, 0, 0, 0], v, None, np.zeros(3)))
    finally:
        lib().orc_soap_free(handle)

    names = ["system", "first_atom", "second_atom", "cell_shift_a", "cell_shift_b", "cell_shift_c"]
    keys, blocks = [], []
    channels = [(l, _radial_size(r)) for l, r in basis_channels(hypers["basis"])]
    for tp in type_pairs:
        for l, nl in channels:
            sl = slice(l * l, (l + 1) * (l + 1))
            entries = [e for e in samples[tp] if l == 0 or not _is_self_pair(e[0])]
            keys.append([l, 1, tp[0], tp[1]])
            values = np.array([e[1][sl, :nl] for e in entries]).reshape(len(entries), 2 * l + 1, nl)
            mu = ("o3_mu", np.arange(-l, l + 1, dtype=np.int32).reshape(-1, 1))
            blk = _block(values, names, [e[0] for e in entries], [mu], ["n"], np.arange(nl).reshape(-1, 1))
            zero = np.zeros((3, 2 * l + 1, nl))
            grad_of = [zero if e[2] is None else e[2][:, sl, :nl] for e in entries]
            if "positions" in gradients:
                rows, gv = [], []
                for k, e in enumerate(entries):
                    system, first, second, sa, sb, sc = e[0]
                    if first == second and (sa, sb, sc) == (0, 0, 0):
                        continue
                    if first == second:
                        rows.append([k, system, first])
                        gv.append(-grad_of[k] + grad_of[k])
                    else:
                        rows += [[k, system, first], [k, system, second]]
                        gv += [-grad_of[k], grad_of[k]]
                gvals = np.array(gv).reshape(len(rows), 3, 2 * l + 1, nl)
                blk["gradients"]["positions"] = _block(gvals, ["sample", "system", "atom"], rows, [_xyz(), mu],
                                                       ["n"], np.arange(nl).reshape(-1, 1))
            for param in ("strain", "cell"):
                if param not in gradients:
                    continue
                gvals = np.zeros((len(entries), 3, 3, 2 * l + 1, nl))
                for k, e in enumerate(entries):
                    w = e[3] if param == "strain" else np.array(e[0][3:6], dtype=np.float64)
                    gvals[k] = w[:, None, None, None] * grad_of[k][None]
                comps = ([("xyz_2", np.arange(3, dtype=np.int32).reshape(3, 1)),
                          ("xyz_1", np.arange(3, dtype=np.int32).reshape(3, 1)), mu] if param == "strain"
                         else [_xyz(), ("abc", np.arange(3, dtype=np.int32).reshape(3, 1)), mu])
                blk["gradients"][param] = _block(gvals, ["sample"], np.arange(len(entries)).reshape(-1, 1), comps,
                                                 ["n"], np.arange(nl).reshape(-1, 1))
            blocks.append(blk)
    return {"keys_names": ["o3_lambda", "o3_sigma", "first_atom_type", "second_atom_type"],
            "keys": np.asarray(keys, dtype=np.int32).reshape(len(keys), 4), "blocks": blocks}


def _is_self_pair(labels):
    return labels[1] == labels[2] and labels[3] == 0 and labels[4] == 0 and labels[5] == 0


def soap_radial_spectrum(systems, hypers, gradients=()):
    """SoapRadialSpectrum (featomic/src/calculators/soap/radial_spectrum.rs:75-135,
    :230-330): the spherical expansion with one angular channel (l = 0), keys
    (center_type, neighbor_type), no components, properties n; gradients keep the
    expansion's rows with the o3_mu component dropped."""
    basis = dict(hypers["basis"])
    se_hypers = {"cutoff": hypers["cutoff"], "density": hypers["density"],
                 "basis": {"type": "TensorProduct", "max_angular": 0, "radial": basis["radial"]}}
    if "spline_accuracy" in basis:
        se_hypers["basis"]["spline_accuracy"] = basis["spline_accuracy"]
    se = soap_spherical_expansion(systems, se_hypers, gradients)
    keys = np.asarray(se["keys"])[:, 2:4]
    blocks = []
    for blk in se["blocks"]:
        ns = len(blk["samples"])
        out = _block(blk["values"].reshape(ns, -1), blk["samples_names"], blk["samples"], [],
                     blk["properties_names"], blk["properties"])
        for param, g in blk["gradients"].items():
            rows = len(g["samples"])
            shape = g["values"].shape  # [rows][gradient components...][1][n]
            values = g["values"].reshape(shape[:-2] + (shape[-1],))
            comps = g["components"][:-1]
            gb = _block(values, g["samples_names"], g["samples"], comps, g["properties_names"], g["properties"])
            out["gradients"][param] = gb
        blocks.append(out)
    return {"keys_names": ["center_type", "neighbor_type"], "keys": keys, "blocks": blocks}


def soap_power_spectrum(systems, hypers, gradients=()):
    comp = SoapComputation(systems, hypers, gradients)
    lmax, N = comp.lmax, comp.N

    # neighbour types around each atom, incl. itself (labels/keys.rs:116-143)
    ntypes = []
    for sd in comp.data:
        t = sd.system.types
        per_atom = []
        for i in range(len(t)):
            s = {int(t[i])}
            for (other, _) in sd.neighbors[i]:
                s.add(int(t[other]))
            per_atom.append(s)
        ntypes.append(per_atom)

    keyset = set()
    for si, sd in enumerate(comp.data):
        t = sd.system.types
        for i in range(len(t)):
            for a in ntypes[si][i]:
                for b in ntypes[si][i]:
                    if b >= a:
                        keyset.add((int(t[i]), a, b))

    # properties (power_spectrum.rs:498-523): (l, n_1, n_2) over the basis' channels;
    # `cols` picks them from the dense [lmax + 1][N][N] layout computed below
    props = [(l, n1, n2) for l, nl in comp.channels for n1 in range(nl) for n2 in range(nl)]
    cols = np.array([l * N * N + n1 * N + n2 for (l, n1, n2) in props], dtype=np.int64)
    dense = (lmax + 1) * N * N
    norms = np.array([(1.0 if l % 2 == 0 else -1.0) * math.sqrt(2 * l + 1) for l in range(lmax + 1)])

    keys, blocks = [], []
    for (ct, t1, t2) in sorted(keyset):
        samples = []
        for si, sd in enumerate(comp.data):
            t = sd.system.types
            for i in range(len(t)):
                if t[i] == ct and t1 in ntypes[si][i] and t2 in ntypes[si][i]:
                    samples.append((si, i))
        kappa = math.sqrt(2.0) if t1 != t2 else 1.0
        i1, i2 = comp.tindex(t1), comp.tindex(t2)
        # power_spectrum.rs:584-618, all samples of the block at once
        c1s = np.array([comp.data[si].values[i, i1] for (si, i) in samples]).reshape(len(samples), comp.L2, N)
        c2s = np.array([comp.data[si].values[i, i2] for (si, i) in samples]).reshape(len(samples), comp.L2, N)
        values = np.zeros((len(samples), dense))
        for l in range(lmax + 1):
            sl = slice(l * l, (l + 1) * (l + 1))
            values[:, l * N * N:(l + 1) * N * N] = (
                kappa * np.einsum("sma,smb->sab", c1s[:, sl], c2s[:, sl]) / norms[l]).reshape(len(samples), -1)
        blk = _block(values[:, cols], ["system", "atom"], samples, [], ["l", "n_1", "n_2"], props)

        if "positions" in comp.gradients:
            grad_samples = []
            for s_i, (si, i) in enumerate(samples):
                sd = comp.data[si]
                t = sd.system.types
                nbrs = set()
                if t[i] in (t1, t2):
                    nbrs.add(i)
                for (other, _) in sd.neighbors[i]:
                    if t[other] in (t1, t2):
                        nbrs.add(other)
                        nbrs.add(i)
                for j in sorted(nbrs):
                    grad_samples.append((s_i, si, j))
            # power_spectrum.rs:621-678; the rows of each system are contracted in C from the
            # table of its spherical-expansion gradient rows (entry 0: no such term)
            gv = np.zeros((len(grad_samples), 3, len(props)))
            r0 = 0
            while r0 < len(grad_samples):
                si = grad_samples[r0][1]
                r1 = r0
                while r1 < len(grad_samples) and grad_samples[r1][1] == si:
                    r1 += 1
                sd = comp.data[si]
                table, self_base, pair_index = comp.se_gradient_table(sd)
                t = sd.system.types
                n = r1 - r0
                center = np.empty(n, dtype=np.int64)
                idx1 = np.zeros(n, dtype=np.int64)
                idx2 = np.zeros(n, dtype=np.int64)
                for r in range(n):
                    s_i, _, j = grad_samples[r0 + r]
                    i = samples[s_i][1]
                    center[r] = i
                    if j == i:
                        idx1[r] = self_base + i * comp.S + i1
                        idx2[r] = self_base + i * comp.S + i2
                    else:
                        # SE gradient sample exists in block t1 iff type_j == t1
                        k = pair_index.get((i, j), 0)
                        idx1[r] = k if t[j] == t1 else 0
                        idx2[r] = k if t[j] == t2 else 0
                out = np.empty((n, 3, dense))
                lib().orc_ps_gradient_rows(n, lmax + 1, N, comp.S, i1, i2, kappa, _ptr(sd.values), _ptr(center),
                                           _ptr(idx1), _ptr(idx2), _ptr(table), _ptr(out))
                gv[r0:r1] = out[:, :, cols]
                r0 = r1
            blk["gradients"]["positions"] = _block(gv, ["sample", "system", "atom"], grad_samples, [_xyz()], ["l", "n_1", "n_2"], props)

        for param, attr in (("cell", "cell_g"), ("strain", "strain")):
            if param not in comp.gradients:
                continue
            gv = np.zeros((len(samples), 3, 3, len(props)))
            for s_i, (si, i) in enumerate(samples):
                sd = comp.data[si]
                c1, c2 = sd.values[i, i1], sd.values[i, i2]
                g1, g2 = getattr(sd, attr)[i, i1], getattr(sd, attr)[i, i2]
                row = []
                for l in range(lmax + 1):
                    sl = slice(l * l, (l + 1) * (l + 1))
                    acc = np.einsum("mb,xyma->xyab", c2[sl], g1[:, :, sl]) + np.einsum("ma,xymb->xyab", c1[sl], g2[:, :, sl])
                    row.append((kappa * acc / norms[l]).reshape(3, 3, -1))
                gv[s_i] = np.concatenate(row, axis=2)[:, :, cols]
            if param == "cell":
                comps = [_xyz(), ("abc", np.arange(3).reshape(3, 1))]
            else:
                comps = [("xyz_2", np.arange(3).reshape(3, 1)), ("xyz_1", np.arange(3).reshape(3, 1))]
            blk["gradients"][param] = _block(gv, ["sample"], np.arange(len(samples)).reshape(-1, 1), comps, ["l", "n_1", "n_2"], props)

        keys.append((ct, t1, t2))
        blocks.append(blk)
    return {"keys_names": ["center_type", "neighbor_1_type", "neighbor_2_type"],
            "keys": np.array(keys, dtype=np.int32).reshape(len(keys), 3), "blocks": blocks}


def soap_ps_baseline(systems, hypers, n_threads=1):
    """Timed CPU baseline: returns (number of PS gradient rows, checksum)."""
    h = soap_hypers(hypers)
    handle = lib().orc_soap_new(ctypes.byref(h))
    species = _species(systems)
    offsets = np.zeros(len(systems) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([len(s) for s in systems])
    positions = np.ascontiguousarray(np.concatenate([s.positions for s in systems]))
    types = np.ascontiguousarray(np.concatenate([s.types for s in systems]))
    cells = np.ascontiguousarray(np.stack([s.cell.reshape(9) for s in systems]))
    checksum = ctypes.c_double()
    rows = lib().orc_soap_ps_batch(handle, len(systems), _ptr(offsets), _ptr(positions), _ptr(types), _ptr(cells),
                                   len(species), _ptr(species), int(n_threads), ctypes.byref(checksum))
    lib().orc_soap_free(handle)
    return rows, checksum.value


# ----------------------------------------------------------------------------
# LODE


def k_vectors(cell, k_cutoff):
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
    n = lib().orc_k_vectors(_ptr(cell), float(k_cutoff), None)
    out = np.zeros((n, 4))
    lib().orc_k_vectors(_ptr(cell), float(k_cutoff), _ptr(out))
    return out


def _lode_terms(hypers, fn):
    h = lode_hypers(hypers)
    handle = lib().orc_lode_new(ctypes.byref(h))
    if not handle:
        raise ValueError("invalid LODE hyper-parameters")
    try:
        return fn(handle, h)
    finally:
        lib().orc_lode_free(handle)


def lode_density_fourier(hypers, k_norms):
    """rho(k) for each |k| (lode/spherical_expansion.rs:280-342)."""
    return _lode_terms(hypers, lambda hd, h: np.array([lib().orc_lode_density_fourier(hd, float(k)) for k in k_norms]))


def lode_center_contribution(hypers):
    """Central-atom term per radial function (lode/radial_integral/gto.rs:136-185)."""
    def fn(hd, h):
        out = np.zeros(h.n_radial)
        lib().orc_lode_center_contribution(hd, _ptr(out))
        return out
    return _lode_terms(hypers, fn)


def lode_k0_contribution(hypers):
    """k = 0 term per radial function (lode/spherical_expansion.rs:347-390)."""
    def fn(hd, h):
        out = np.zeros(h.n_radial)
        lib().orc_lode_k0_contribution(hd, _ptr(out))
        return out
    return _lode_terms(hypers, fn)


def lode_dense(system, hypers, species, centers=()):
    """Dense LODE expansion of one system: values [n][S][L2][N] of every atom and the
    position-gradient rows of the listed centers, grads [c][n][3][L2][N] (row (i, j) =
    d c_i^{type j} / d r_j) and self_grads [c][S][3][L2][N]."""
    h = lode_hypers(hypers)
    handle = lib().orc_lode_new(ctypes.byref(h))
    if not handle:
        raise ValueError("invalid LODE hyper-parameters")
    L2, N = (h.max_angular + 1) ** 2, h.n_radial
    n, S = len(system), len(species)
    species = np.ascontiguousarray(species, dtype=np.int32)
    centers = np.ascontiguousarray(centers, dtype=np.int32)
    values = np.zeros((n, S, L2, N))
    grads = np.zeros((len(centers), n, 3, L2, N))
    self_grads = np.zeros((len(centers), S, 3, L2, N))
    try:
        status = lib().orc_lode_system_centers(handle, n, _ptr(system.positions), _ptr(system.types),
                                               _ptr(system.cell.reshape(9)), S, _ptr(species), _ptr(values),
                                               len(centers), _ptr(centers) if len(centers) else None,
                                               _ptr(grads) if len(centers) else None,
                                               _ptr(self_grads) if len(centers) else None)
    finally:
        lib().orc_lode_free(handle)
    if status != 0:
        raise ValueError(f"LODE oracle failed with status {status}")
    return values, grads, self_grads


def lode_spherical_expansion(systems, hypers, gradients=()):
    h = lode_hypers(hypers)
    handle = lib().orc_lode_new(ctypes.byref(h))
    if not handle:
        raise ValueError("invalid LODE hyper-parameters")
    lmax, N = h.max_angular, h.n_radial
    L2 = (lmax + 1) ** 2
    species = _species(systems)
    S = len(species)
    do_grad = "positions" in gradients
    dense = []
    try:
        for system in systems:
            n = len(system)
            values = np.zeros((n, S, L2, N))
            grads = np.zeros((n, n, 3, L2, N)) if do_grad else None
            self_grads = np.zeros((n, S, 3, L2, N)) if do_grad else None
            status = lib().orc_lode_system(handle, n, _ptr(system.positions), _ptr(system.types), _ptr(system.cell.reshape(9)),
                                           S, _ptr(species), _ptr(values), _ptr(grads), _ptr(self_grads))
            if status != 0:
                raise ValueError(f"LODE oracle failed with status {status}")
            dense.append((values, grads, self_grads))
    finally:
        lib().orc_lode_free(handle)

    # AllTypesPairsKeys (labels/keys.rs:36-58)
    keyset = set()
    for s in systems:
        for a in s.types:
            for b in s.types:
                keyset.add((int(a), int(b)))
    keys, blocks = [], []
    for (ct, nt) in sorted(keyset):
        samples = []
        for si, s in enumerate(systems):
            if np.any(s.types == nt):
                for i in range(len(s)):
                    if s.types[i] == ct:
                        samples.append((si, i))
        grad_samples = []
        for s_i, (si, i) in enumerate(samples):
            for j in range(len(systems[si])):
                if systems[si].types[j] == nt or j == i:
                    grad_samples.append((s_i, si, j))
        nti = int(np.searchsorted(species, nt))
        for l in range(lmax + 1):
            sl = slice(l * l, (l + 1) * (l + 1))
            keys.append((l, 1, ct, nt))
            values = np.zeros((len(samples), 2 * l + 1, N))
            for s_i, (si, i) in enumerate(samples):
                values[s_i] = dense[si][0][i, nti, sl]
            mu = ("o3_mu", np.arange(-l, l + 1, dtype=np.int32).reshape(-1, 1))
            blk = _block(values, ["system", "atom"], samples, [mu], ["n"], np.arange(N).reshape(-1, 1))
            if do_grad:
                gv = np.zeros((len(grad_samples), 3, 2 * l + 1, N))
                for r, (s_i, si, j) in enumerate(grad_samples):
                    i = samples[s_i][1]
                    if i == j:
                        gv[r] = dense[si][2][i, nti, :, sl]
                    else:
                        gv[r] = dense[si][1][i, j, :, sl]
                blk["gradients"]["positions"] = _block(gv, ["sample", "system", "atom"], grad_samples, [_xyz(), mu], ["n"], np.arange(N).reshape(-1, 1))
            blocks.append(blk)
    return {"keys_names": ["o3_lambda", "o3_sigma", "center_type", "neighbor_type"],
            "keys": np.array(keys, dtype=np.int32).reshape(len(keys), 4), "blocks": blocks}
