"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/*.h, and the host-side parts of the API (construction,
hyper-parameter validation, name / parameters / cutoffs, buffer sizes, error
codes) behave like featomic/tests/c/calculator.cpp and
python/featomic/tests/calculators/hypers.py expect. Nothing here launches a
kernel."""

import ctypes
import json
import os
import re

import numpy as np
import pytest

from rascaline_amd import _c_lib as C
from rascaline_amd import calculators

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

FEATOMIC_SUCCESS = 0
FEATOMIC_INVALID_PARAMETER_ERROR = 1
FEATOMIC_BUFFER_SIZE_ERROR = 254

SOAP = {
    "cutoff": {"radius": 5.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 6, "radial": {"type": "Gto", "max_radial": 7}},
}
LODE = {
    "density": {"type": "SmearedPowerLaw", "smearing": 1.0, "exponent": 1},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5, "radius": 5.0}},
}


def declared_functions():
    names = set()
    pattern = re.compile(r"^[A-Za-z_][A-Za-z0-9_ \*]*[ \*]((?:featomic|mts)[a-z0-9_]*)\(", re.M)
    for header in ("featomic.h", "metatensor.h", "featomic_amd.h"):
        with open(os.path.join(ROOT, "include", header)) as fd:
            names.update(pattern.findall(fd.read()))
    return sorted(names)


def make(name, hypers):
    text = hypers if isinstance(hypers, str) else json.dumps(hypers)
    return C.lib().featomic_calculator(name.encode(), text.encode())


def last_error():
    return C.lib().featomic_last_error().decode()


def test_every_declared_function_is_exported():
    names = declared_functions()
    # the reference API (featomic.h), the metatensor subset and the extensions
    assert len(names) >= 25
    handle = ctypes.CDLL(C.LIBRARY_PATH)
    missing = [name for name in names if not hasattr(handle, name)]
    assert missing == []
    # and the Python prototypes cover all of them
    assert {p[0] for p in C.PROTOTYPES} >= set(names)


@pytest.mark.parametrize(
    "name, hypers, expected",
    [
        # names from soap/power_spectrum.rs:412, soap/spherical_expansion.rs:651,
        # lode/spherical_expansion.rs:454
        ("soap_power_spectrum", SOAP, "SOAP power spectrum"),
        ("spherical_expansion", SOAP, "spherical expansion"),
        ("lode_spherical_expansion", LODE, "lode spherical expansion"),
        # soap/spherical_expansion_pair.rs:548-550
        ("spherical_expansion_by_pair", SOAP, "spherical expansion by pair"),
    ],
)
def test_name_parameters_cutoffs(name, hypers, expected):
    text = json.dumps(hypers)
    calc = make(name, text)
    assert calc, last_error()
    try:
        buffer = ctypes.create_string_buffer(256)
        assert C.lib().featomic_calculator_name(calc, buffer, 256) == FEATOMIC_SUCCESS
        assert buffer.value.decode() == expected

        buffer = ctypes.create_string_buffer(4096)
        assert C.lib().featomic_calculator_parameters(calc, buffer, 4096) == FEATOMIC_SUCCESS
        assert buffer.value.decode() == text

        # too small a buffer: FEATOMIC_BUFFER_SIZE_ERROR (featomic/tests/c/calculator.cpp)
        small = ctypes.create_string_buffer(8)
        assert C.lib().featomic_calculator_parameters(calc, small, 8) == FEATOMIC_BUFFER_SIZE_ERROR

        cutoffs = ctypes.POINTER(ctypes.c_double)()
        count = ctypes.c_size_t()
        assert C.lib().featomic_calculator_cutoffs(calc, ctypes.byref(cutoffs), ctypes.byref(count)) == FEATOMIC_SUCCESS
        values = [cutoffs[k] for k in range(count.value)]
        if name == "lode_spherical_expansion":
            assert values == []
        else:
            assert values == [5.0]
    finally:
        C.lib().featomic_calculator_free(calc)


def test_unknown_calculator():
    calc = make("not_a_calculator", "{}")
    assert not calc
    assert "not_a_calculator" in last_error()


@pytest.mark.parametrize(
    "hypers, prefix",
    [
        ("{", "json error"),
        (dict(SOAP, extra=1), "json error"),
        ({k: v for k, v in SOAP.items() if k != "density"}, "json error"),
        (dict(SOAP, cutoff={"radius": -1.0, "smoothing": {"type": "Step"}}), "invalid parameter"),
        (dict(SOAP, density={"type": "Gaussian", "width": -0.3}), "invalid parameter"),
    ],
)
def test_invalid_soap_hypers(hypers, prefix):
    """Errors surface as a null calculator and a message in featomic_last_error
    whose prefix names the error kind (featomic/src/errors.rs Display); the
    Python layer raises FeatomicError with that message."""
    calc = make("spherical_expansion", hypers)
    assert not calc
    assert last_error().startswith(prefix), last_error()
    text = hypers if isinstance(hypers, str) else json.dumps(hypers)
    with pytest.raises(C.FeatomicError, match=prefix):
        calculators.CalculatorBase("spherical_expansion", text)


def test_null_arguments():
    calc = make("soap_power_spectrum", SOAP)
    try:
        assert C.lib().featomic_calculator_name(calc, None, 10) == FEATOMIC_INVALID_PARAMETER_ERROR
        assert "null" in last_error().lower()
    finally:
        C.lib().featomic_calculator_free(calc)
    # freeing NULL is allowed
    C.lib().featomic_calculator_free(None)


def test_precision_extension():
    calc = make("soap_power_spectrum", SOAP)
    try:
        assert C.lib().featomic_amd_calculator_set_precision(calc, 32) == FEATOMIC_SUCCESS
        assert C.lib().featomic_amd_calculator_set_precision(calc, 64) == FEATOMIC_SUCCESS
        assert C.lib().featomic_amd_calculator_set_precision(calc, 16) == FEATOMIC_INVALID_PARAMETER_ERROR
    finally:
        C.lib().featomic_calculator_free(calc)


def test_profiling_and_logging_entry_points():
    assert C.lib().featomic_profiling_clear() == FEATOMIC_SUCCESS
    assert C.lib().featomic_profiling_enable(True) == FEATOMIC_SUCCESS
    assert C.lib().featomic_profiling_enable(False) == FEATOMIC_SUCCESS
    buffer = ctypes.create_string_buffer(4096)
    assert C.lib().featomic_profiling_get(b"short_table", buffer, 4096) == FEATOMIC_SUCCESS
    assert C.lib().featomic_profiling_get(b"json", buffer, 4096) == FEATOMIC_SUCCESS
    json.loads(buffer.value.decode())
    assert C.lib().featomic_profiling_get(b"no-such-format", buffer, 4096) == FEATOMIC_INVALID_PARAMETER_ERROR


def test_radial_surrogate_host_checks():
    opts = C.featomic_amd_surrogate_options_t()
    assert C.lib().featomic_amd_surrogate_options_default(ctypes.byref(opts)) == FEATOMIC_SUCCESS
    assert (opts.n_samples, opts.steps, opts.batch, opts.n_validation) == (1000000, 20000, 2048, 20001)
    assert (opts.learning_rate, opts.ridge, opts.tolerance, opts.seed) == (3e-3, 1e-12, 1e-3, 1)
    calc = make("soap_power_spectrum", SOAP)
    try:
        # nothing trained yet: the spline stays in use
        assert C.lib().featomic_amd_calculator_use_radial_surrogate(calc, 1) == FEATOMIC_INVALID_PARAMETER_ERROR
        assert "no accepted radial surrogate" in last_error()
        assert C.lib().featomic_amd_calculator_use_radial_surrogate(calc, 0) == FEATOMIC_SUCCESS
        assert C.lib().featomic_amd_calculator_use_radial_surrogate(calc, 2) == FEATOMIC_INVALID_PARAMETER_ERROR
        required = ctypes.c_int64()
        assert C.lib().featomic_amd_calculator_radial_surrogate_parameters(calc, None, 0, ctypes.byref(required)) \
            == FEATOMIC_INVALID_PARAMETER_ERROR
        bad = C.featomic_amd_surrogate_options_t()
        C.lib().featomic_amd_surrogate_options_default(ctypes.byref(bad))
        bad.n_samples = 0
        assert C.lib().featomic_amd_calculator_train_radial_surrogate(calc, ctypes.byref(bad), None) \
            == FEATOMIC_INVALID_PARAMETER_ERROR
        assert "invalid radial surrogate options" in last_error()
    finally:
        C.lib().featomic_calculator_free(calc)
    lode = make("lode_spherical_expansion", json.dumps({
        "density": {"type": "SmearedPowerLaw", "smearing": 1.0, "exponent": 1},
        "basis": {"type": "TensorProduct", "max_angular": 2, "radial": {"type": "Gto", "max_radial": 2, "radius": 3.0}}}))
    try:
        assert C.lib().featomic_amd_calculator_train_radial_surrogate(lode, ctypes.byref(opts), None) \
            == FEATOMIC_INVALID_PARAMETER_ERROR
        assert "only available for the SOAP calculators" in last_error()
    finally:
        C.lib().featomic_calculator_free(lode)
    # network kernels reject sizes before touching the GPU
    loss = ctypes.c_double()
    buf = np.zeros(4, dtype=np.float32)
    ptr = buf.ctypes.data_as(ctypes.c_void_p)
    assert C.lib().featomic_amd_mlp_loss_gradient(0, 1, ptr, ptr, ptr, ctypes.byref(loss), ptr) \
        == FEATOMIC_INVALID_PARAMETER_ERROR
    assert C.lib().featomic_amd_mlp_forward(513, 1, ptr, ptr, ptr) == FEATOMIC_INVALID_PARAMETER_ERROR


def test_tuning_host_checks():
    calc = make("soap_power_spectrum", SOAP)
    try:
        e, w = ctypes.c_int32(-5), ctypes.c_int32(-5)
        assert C.lib().featomic_amd_calculator_get_tuning(calc, ctypes.byref(e), ctypes.byref(w)) == FEATOMIC_SUCCESS
        assert (e.value, w.value) == (0, -1)
        assert C.lib().featomic_amd_calculator_set_tuning(calc, 16, 3) == FEATOMIC_SUCCESS
        C.lib().featomic_amd_calculator_get_tuning(calc, ctypes.byref(e), ctypes.byref(w))
        assert (e.value, w.value) == (16, 3)
        for entries, window in [(12, -1), (128, -1), (-1, -1), (8, -3), (8, 1000)]:
            assert C.lib().featomic_amd_calculator_set_tuning(calc, entries, window) == FEATOMIC_INVALID_PARAMETER_ERROR
        # the rejected values left the last accepted ones in place
        C.lib().featomic_amd_calculator_get_tuning(calc, ctypes.byref(e), ctypes.byref(w))
        assert (e.value, w.value) == (16, 3)
        assert C.lib().featomic_amd_calculator_autotune(calc, None, None, 0, None, 3, None, 0) \
            == FEATOMIC_INVALID_PARAMETER_ERROR
    finally:
        C.lib().featomic_calculator_free(calc)


def test_positions_vjp_host_checks():
    calc = make("soap_power_spectrum", SOAP)
    try:
        assert C.lib().featomic_amd_positions_vjp(calc, None, None, None, None) == FEATOMIC_INVALID_PARAMETER_ERROR
        assert "systems" in last_error()
        systems = C.featomic_amd_device_systems_t()
        offsets = np.zeros(1, dtype=np.int64)
        cells = np.zeros(9)
        systems.n_systems = 0
        systems.atom_offsets = offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        systems.cells = cells.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        assert C.lib().featomic_amd_positions_vjp(calc, ctypes.byref(systems), None, None, None) \
            == FEATOMIC_INVALID_PARAMETER_ERROR
        assert "grad_values" in last_error()
    finally:
        C.lib().featomic_calculator_free(calc)


def test_radial_spectrum_parameters():
    radial = {"cutoff": {"radius": 4.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
              "density": {"type": "Gaussian", "width": 0.3},
              "basis": {"radial": {"type": "Gto", "max_radial": 3}}}
    calc = make("soap_radial_spectrum", json.dumps(radial))
    try:
        buf = ctypes.create_string_buffer(64)
        assert C.lib().featomic_calculator_name(calc, buf, 64) == FEATOMIC_SUCCESS
        assert buf.value == b"radial spectrum"
    finally:
        C.lib().featomic_calculator_free(calc)
    # the basis of the radial spectrum has no type or max_angular (deny_unknown_fields)
    bad = dict(radial, basis={"type": "TensorProduct", "max_angular": 2, "radial": {"type": "Gto", "max_radial": 3}})
    assert not C.lib().featomic_calculator(b"soap_radial_spectrum", json.dumps(bad).encode())
    assert "unknown field" in last_error()


def test_neighbor_list_parameters():
    """NeighborList parameters (featomic/src/calculators/neighbor_list.rs:31-44): three
    required fields, unknown fields ignored (no deny_unknown_fields); name "neighbors
    list" (neighbor_list.rs:56-58); one cutoff."""
    params = {"cutoff": 3.5, "full_neighbor_list": True, "self_pairs": False}
    calc = make("neighbor_list", dict(params, extra="ignored"))
    assert calc, last_error()
    try:
        buf = ctypes.create_string_buffer(64)
        assert C.lib().featomic_calculator_name(calc, buf, 64) == FEATOMIC_SUCCESS
        assert buf.value == b"neighbors list"
        cutoffs = ctypes.POINTER(ctypes.c_double)()
        count = ctypes.c_size_t()
        assert C.lib().featomic_calculator_cutoffs(calc, ctypes.byref(cutoffs), ctypes.byref(count)) == FEATOMIC_SUCCESS
        assert [cutoffs[k] for k in range(count.value)] == [3.5]
    finally:
        C.lib().featomic_calculator_free(calc)
    for missing in params:
        assert not make("neighbor_list", {k: v for k, v in params.items() if k != missing})
        assert last_error() == f"json error: missing field `{missing}`"
    assert not make("neighbor_list", dict(params, self_pairs=1))
    assert last_error().startswith("json error: invalid type for `self_pairs`")
    nl = calculators.NeighborList(cutoff=2.0, full_neighbor_list=False)
    assert json.loads(nl.parameters) == {"cutoff": 2.0, "full_neighbor_list": False, "self_pairs": False}


def test_neighbor_list_rejects_cell_gradients():
    """Only positions gradients (neighbor_list.rs:100-106); the check runs before
    any system is read (calculator.rs:342-380)."""
    import rascaline_amd as ra

    nl = ra.NeighborList(cutoff=2.0, full_neighbor_list=False)
    system = ra.System(np.array([1, 1]), np.zeros((2, 3)), np.eye(3) * 5.0)
    for grad in ("cell", "strain"):
        with pytest.raises(C.FeatomicError, match=f"the neighbors list calculator does not support gradients with respect to {grad}"):
            nl.compute(system, gradients=[grad])


def test_by_pair_parameters_deny_unknown_fields():
    """SphericalExpansionByPair takes the expansion's parameters, which deny unknown
    fields (soap/spherical_expansion_pair.rs:37-60)."""
    assert not make("spherical_expansion_by_pair", dict(SOAP, extra=1))
    assert last_error().startswith("json error: unknown field")


@pytest.mark.parametrize("name, hypers", [("neighbor_list", {"cutoff": 3.0, "full_neighbor_list": True, "self_pairs": False}),
                                          ("spherical_expansion_by_pair", SOAP)])
def test_pair_calculators_have_no_device_path(name, hypers):
    """The pair calculators return host TensorMaps only: the device-resident entry
    point rejects them before touching any device memory."""
    calc = make(name, hypers)
    assert calc, last_error()
    try:
        systems = C.featomic_amd_device_systems_t()
        offsets = np.zeros(1, dtype=np.int64)
        cells = np.zeros(9)
        systems.n_systems = 0
        systems.atom_offsets = offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        systems.cells = cells.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        output = C.featomic_amd_device_output_t()
        status = C.lib().featomic_amd_compute_device(calc, ctypes.byref(systems), None, 0, None, ctypes.byref(output))
        assert status == FEATOMIC_INVALID_PARAMETER_ERROR
        assert "host TensorMaps only" in last_error()
    finally:
        C.lib().featomic_calculator_free(calc)


def _points(n_points, size, start=0.0, stop=4.0):
    return [{"position": start + (stop - start) * k / max(n_points - 1, 1), "values": [0.1 * k] * size,
             "derivatives": [0.0] * size} for k in range(n_points)]


TABULATED = {"type": "Tabulated", "points": _points(5, 3)}


@pytest.mark.parametrize(
    "density, basis, message",
    [
        # (density, radial basis) pairs of soap/radial_integral/mod.rs:85-113
        ({"type": "DiracDelta"}, {"type": "TensorProduct", "max_angular": 2, "radial": {"type": "Gto", "max_radial": 3}},
         "this combination of basis and density is not supported in SOAP"),
        ({"type": "Gaussian", "width": 0.3}, {"type": "TensorProduct", "max_angular": 2, "radial": TABULATED},
         "this combination of basis and density is not supported in SOAP"),
        # tabulated.rs:171-208
        ({"type": "DiracDelta"}, {"type": "TensorProduct", "max_angular": 2,
                                  "radial": {"type": "Tabulated", "points": _points(1, 3)}},
         "we need at least two points to define a 'Tabulated' radial basis"),
        ({"type": "DiracDelta"}, {"type": "TensorProduct", "max_angular": 2,
                                  "radial": {"type": "Tabulated", "points": _points(3, 3) + _points(2, 2)}},
         "expected all points 'values' in 'Tabulated' radial basis to have the same size"),
        ({"type": "DiracDelta"}, {"type": "Explicit", "by_angular": {"x": TABULATED}}, "json error"),
        ({"type": "DiracDelta"}, {"type": "Explicit", "by_angular": {}}, "at least one angular channel"),
        ({"type": "Gaussian", "width": 0.3},
         {"type": "Explicit", "by_angular": {"0": {"type": "Gto", "max_radial": 3}, "2": {"type": "Gto", "max_radial": 5}},
          "spline_accuracy": None},
         "needs spline_accuracy"),
    ],
)
def test_invalid_basis_and_density(density, basis, message):
    hypers = dict(SOAP, density=density, basis=basis)
    calc = make("spherical_expansion", hypers)
    assert not calc
    assert message in last_error(), last_error()


@pytest.mark.parametrize("name", ["spherical_expansion", "soap_power_spectrum", "spherical_expansion_by_pair"])
def test_explicit_and_tabulated_bases_are_accepted(name):
    """DiracDelta + Tabulated (TensorProduct or Explicit) and Explicit GTO bases build
    calculators; their parameters round-trip through featomic_calculator_parameters."""
    for density, basis in (
        ({"type": "DiracDelta"}, {"type": "TensorProduct", "max_angular": 3, "radial": TABULATED}),
        ({"type": "DiracDelta", "center_atom_weight": 0.5}, {"type": "Explicit", "by_angular": {"0": TABULATED, "3": TABULATED}}),
        ({"type": "Gaussian", "width": 0.3},
         {"type": "Explicit", "by_angular": {"1": {"type": "Gto", "max_radial": 3}, "0": {"type": "Gto", "max_radial": 5}}}),
    ):
        hypers = dict(SOAP, density=density, basis=basis)
        calc = make(name, hypers)
        assert calc, last_error()
        C.lib().featomic_calculator_free(calc)
