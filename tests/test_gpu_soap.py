"""GPU parity of the SOAP path (neighbour list, spherical expansion, power
spectrum, gradients) against the FP64 oracle and the reference's fixtures.

Tolerances: FP64 kernels within 1e-10 relative, |a-b| <= 1e-10 (|b| + max|b|), where
max|b| is the largest element of the same kind in the whole descriptor (elements that
cancel to ~1e-12 of the scale are compared on the scale); FP32 kernels within 2e-4
in the same sense. Fixture comparisons use the reference's own tolerances.
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import (assert_relative_eq, block_relative_error, compare_tensormaps, flatten_ps, flatten_se, golden,
                     load_npy_gz, oracle_to_tensormap, random_system, sum_gradients)

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def hypers(rc=5.0, lmax=6, nmax=7, width=0.3, smoothing=0.5, accuracy=1e-8, scaling=None):
    h = {
        "cutoff": {"radius": rc, "smoothing": {"type": "ShiftedCosine", "width": smoothing}},
        "density": {"type": "Gaussian", "width": width},
        "basis": {"type": "TensorProduct", "max_angular": lmax, "radial": {"type": "Gto", "max_radial": nmax},
                  "spline_accuracy": accuracy},
    }
    if scaling:
        h["density"]["scaling"] = scaling
    return h


def ra_systems(systems):
    return [ra.System(s.types, s.positions, s.cell) for s in systems]


def orc_systems(raw):
    return [orc.System(p, t, c) for (p, t, c) in raw]


# ---------------------------------------------------------------------------- neighbour list

NL_CASES = [("cluster", 30, 6.0, 3.42), ("periodic", 40, 7.0, 4.5), ("triclinic", 35, 6.0, 5.1),
            ("outside", 25, 6.5, 3.3), ("periodic", 6, 3.0, 5.0), ("triclinic", 5, 2.6, 4.0)]


@pytest.mark.parametrize("case", range(len(NL_CASES)))
def test_neighbors_bit_exact(case):
    kind, n, L, rc = NL_CASES[case]
    say("NL case", NL_CASES[case])
    rng = np.random.default_rng(100 + case)
    pos, types, cell = random_system(rng, n, kind, L)
    got = ra.neighbors(pos, cell, rc)
    exp = orc.neighbors(pos, cell, rc)
    assert len(got) == len(exp)
    assert np.array_equal(got["first"], exp["first"])
    assert np.array_equal(got["second"], exp["second"])
    assert np.array_equal(got["cell_shift_indices"], exp["shift"])
    assert np.array_equal(got["vector"].view(np.uint64), exp["vector"].view(np.uint64))
    assert np.array_equal(got["distance"].view(np.uint64), exp["distance"].view(np.uint64))


def test_neighbors_reference_unit_cases():
    # featomic/src/systems/neighbors.rs:446-561
    cell = np.array([[0.0, 1.5, 1.5], [1.5, 0.0, 1.5], [1.5, 1.5, 0.0]])
    pairs = ra.neighbors(np.zeros((1, 3)), cell, 3.0)
    assert len(pairs) == 6
    assert np.allclose(pairs["distance"], 2.1213203435596424, rtol=4e-16)
    pairs = ra.neighbors(np.zeros((1, 3)), np.eye(3) * 0.5, 0.6)
    assert sorted(tuple(s) for s in pairs["cell_shift_indices"]) == [(0, 0, 1), (0, 1, 0), (1, 0, 0)]
    cell = np.array([[4.26, -2.45951215, 0.0], [2.13, 1.22975607, 0.0], [0.0, 0.0, 50.0]])
    pos = np.array([[1.42, 0, 0], [2.84, 0, 0], [3.55, -1.22975607, 0], [4.97, -1.22975607, 0]])
    pairs = ra.neighbors(pos, cell, 6.4)
    assert len(pairs) == 90
    found = {(int(p["first"]), int(p["second"]), tuple(int(x) for x in p["cell_shift_indices"])) for p in pairs}
    for missing in [(0, 3, (-2, 0, 0)), (0, 3, (-2, 1, 0)), (0, 3, (-2, 2, 0))]:
        assert missing in found


# ---------------------------------------------------------------------------- fixtures


def test_se_fixture_values_exact_radial():
    systems, h = orc.load_input(golden("spherical-expansion-values-input.json"))
    calc = ra.SphericalExpansion(**h)
    tm = calc.compute(ra_systems(systems))
    expected = load_npy_gz("spherical-expansion-values.npy.gz")
    # reference tolerance 5e-5 (soap-spherical-expansion.rs:32) with an absolute floor at
    # 1e-9 of the largest element: the fixture predates the 30-term asymptotic 1F1 series
    assert_relative_eq(flatten_se(tm).block().values, expected, 5e-5, 1e-9 * np.abs(expected).max())
    oracle = oracle_to_tensormap(orc.soap_spherical_expansion(systems, h))
    compare_tensormaps(tm, oracle, 1e-10, 1e-10)


def test_se_fixture_pbc_values():
    systems, h = orc.load_input(golden("spherical-expansion-pbc-values-input.json"))
    tm = ra.SphericalExpansion(**h).compute(ra_systems(systems))
    assert_relative_eq(flatten_se(tm).block().values, load_npy_gz("spherical-expansion-pbc-values.npy.gz"), 1e-8, 1e-12)


def test_se_fixture_gradients():
    systems, h = orc.load_input(golden("spherical-expansion-gradients-input.json"))
    n_atoms = sum(len(s) for s in systems)
    tm = ra.SphericalExpansion(**h).compute(ra_systems(systems), gradients=["positions", "strain", "cell"])
    block = flatten_se(tm).block()
    # reference tolerance 1e-6 (soap-spherical-expansion.rs:89-99)
    assert_relative_eq(sum_gradients(n_atoms, block.gradient("positions")),
                       load_npy_gz("spherical-expansion-positions-gradient.npy.gz"), 1e-6, 1e-12)
    assert_relative_eq(block.gradient("strain").values, load_npy_gz("spherical-expansion-strain-gradient.npy.gz"), 1e-6, 1e-12)
    assert_relative_eq(block.gradient("cell").values, load_npy_gz("spherical-expansion-cell-gradient.npy.gz"), 1e-6, 1e-12)


def test_ps_fixture_values():
    systems, h = orc.load_input(golden("soap-power-spectrum-values-input.json"))
    tm = ra.SoapPowerSpectrum(**h).compute(ra_systems(systems))
    assert_relative_eq(flatten_ps(tm).block().values, load_npy_gz("soap-power-spectrum-values.npy.gz"), 1e-5, 1e-12)


def test_ps_fixture_gradients():
    systems, h = orc.load_input(golden("soap-power-spectrum-gradients-input.json"))
    n_atoms = sum(len(s) for s in systems)
    tm = ra.SoapPowerSpectrum(**h).compute(ra_systems(systems), gradients=["positions", "strain", "cell"])
    block = flatten_ps(tm).block()
    # reference tolerance 1e-6 (soap-power-spectrum.rs:55-65)
    assert_relative_eq(sum_gradients(n_atoms, block.gradient("positions")),
                       load_npy_gz("soap-power-spectrum-positions-gradient.npy.gz"), 1e-6, 1e-12)
    assert_relative_eq(block.gradient("strain").values, load_npy_gz("soap-power-spectrum-strain-gradient.npy.gz"), 1e-6, 1e-12)
    assert_relative_eq(block.gradient("cell").values, load_npy_gz("soap-power-spectrum-cell-gradient.npy.gz"), 1e-6, 1e-12)


# ---------------------------------------------------------------------------- random systems vs oracle

RANDOM_CASES = [
    ("periodic", [(18, 6.0), (25, 7.0)], hypers(rc=4.5, lmax=4, nmax=5)),
    ("cluster", [(20, 6.0), (12, 5.0)], hypers(rc=5.0, lmax=6, nmax=7)),
    ("triclinic", [(14, 5.5)], hypers(rc=5.5, lmax=3, nmax=3, scaling={"type": "Willatt2018", "scale": 1.5, "rate": 0.8, "exponent": 2.0})),
    ("periodic", [(4, 3.2)], hypers(rc=4.0, lmax=2, nmax=2, accuracy=None)),  # many self images
]


def _random(case, seed):
    kind, specs, h = RANDOM_CASES[case]
    rng = np.random.default_rng(seed)
    raw = [random_system(rng, n, kind, L, types=(1, 6, 8)) for (n, L) in specs]
    return orc_systems(raw), h


@pytest.mark.parametrize("case", range(len(RANDOM_CASES)))
def test_se_vs_oracle_fp64(case):
    say("SE case", case)
    systems, h = _random(case, 7 + case)
    grads = ["positions", "strain", "cell"]
    tm = ra.SphericalExpansion(**h).compute(ra_systems(systems), gradients=grads)
    oracle = oracle_to_tensormap(orc.soap_spherical_expansion(systems, h, grads))
    compare_tensormaps(tm, oracle, 1e-10, 1e-10, gradients=grads)


@pytest.mark.parametrize("case", range(len(RANDOM_CASES)))
def test_ps_vs_oracle_fp64(case):
    say("PS case", case)
    systems, h = _random(case, 11 + case)
    grads = ["positions", "strain", "cell"]
    tm = ra.SoapPowerSpectrum(**h).compute(ra_systems(systems), gradients=grads)
    oracle = oracle_to_tensormap(orc.soap_power_spectrum(systems, h, grads))
    compare_tensormaps(tm, oracle, 1e-10, 1e-10, gradients=grads)


@pytest.mark.parametrize("case", [0, 1])
def test_ps_vs_oracle_fp32(case):
    say("PS fp32 case", case)
    systems, h = _random(case, 21 + case)
    calc = ra.SoapPowerSpectrum(**h)
    calc.set_precision(32)
    tm = calc.compute(ra_systems(systems), gradients=["positions"])
    oracle = oracle_to_tensormap(orc.soap_power_spectrum(systems, h, ["positions"]))
    compare_tensormaps(tm, oracle, 2e-4, 2e-4, gradients=["positions"])


def test_user_provided_neighbors():
    # use_native_system=False: the library uses the caller's half neighbour list
    systems, h = _random(0, 31)
    ra_sys = []
    for s in systems:
        pairs = ra.neighbors(s.positions, s.cell, h["cutoff"]["radius"])
        ra_sys.append(ra.System(s.types, s.positions, s.cell, pairs=pairs))
    calc = ra.SoapPowerSpectrum(**h)
    a = calc.compute(ra_sys, gradients=["positions"], use_native_system=False)
    b = calc.compute(ra_systems(systems), gradients=["positions"], use_native_system=True)
    compare_tensormaps(a, b, 1e-13, 1e-14, gradients=["positions"])


# ---------------------------------------------------------------------------- fast form

@pytest.mark.parametrize("kind", ["periodic", "triclinic", "cluster"])
def test_ps_fast_form_vs_oracle(kind):
    """Fast form (bfloat16 hi + lo MFMA contraction, FP32 elsewhere) against the FP64 oracle
    on the benchmark's hyper-parameters (n = 8, l <= 6): the metric's bar, 1e-4 relative
    error per block (max|a - b| <= 1e-4 max|b| in every block, values and gradients)."""
    rng = np.random.default_rng({"periodic": 71, "triclinic": 72, "cluster": 73}[kind])
    systems = []
    for n, L in [(24, 6.2), (40, 7.0)]:
        pos, types, cell = random_system(rng, n, kind, L if kind != "cluster" else 9.0, types=(1, 6, 7, 8))
        systems.append(orc.System(pos, types, cell))
    h = hypers(rc=5.0, lmax=6, nmax=7)
    oracle = oracle_to_tensormap(orc.soap_power_spectrum(systems, h, ["positions"]))
    errors = {}
    for fast in (False, True):
        calc = ra.SoapPowerSpectrum(**h)
        calc.set_precision(32)
        calc.set_fast_form(fast)
        tm = calc.compute(ra_systems(systems), gradients=["positions"])
        compare_tensormaps(tm, oracle, 2e-4, 2e-4, gradients=["positions"])
        errors[fast] = (block_relative_error(tm, oracle), block_relative_error(tm, oracle, "positions"))
    say("block-relative error (values, gradients): plain", errors[False], "fast", errors[True])
    assert max(errors[True]) <= 1e-4, errors[True]
    assert max(errors[False]) <= 1e-4, errors[False]


def per_l_pair_errors(actual, expected, L, n, param=None):
    """max|a - b| / max|b| for every (l, block) slice: the error per (l, type pair)."""
    get = (lambda b: b.values) if param is None else (lambda b: b.gradient(param).values)
    out = {}
    for key, ba, be in zip(expected.keys.values, actual.blocks(), expected.blocks()):
        a = np.asarray(get(ba)).reshape(-1, L, n * n)
        b = np.asarray(get(be)).reshape(-1, L, n * n)
        if not b.size:
            continue
        for l in range(L):
            scale = float(np.abs(b[:, l]).max())
            if scale > 0:
                out[(l, tuple(int(k) for k in key))] = float(np.abs(a[:, l] - b[:, l]).max()) / scale
    return out


@pytest.mark.parametrize("nmax,lmax,n_atoms", [(11, 8, 36), (15, 4, 30)])
def test_ps_fast_form_bulk_per_l_and_type_pair(nmax, lmax, n_atoms):
    """Mixed-precision power spectrum of the bulk hyper-parameters (rc 6 A, 12 radial
    functions, l <= 8, 5 types uniform, 0.1 atoms/A^3; and n = 16 at l <= 4) against
    the FP64 oracle: the error of every (l, center type, neighbour type pair) slice,
    max|a - b| / max|b|, is within 1e-4 for values and position gradients."""
    say("bulk fast form, n", nmax + 1, "l <=", lmax, "atoms", n_atoms)
    rng = np.random.default_rng(97 + nmax)
    side = (n_atoms / 0.1) ** (1.0 / 3.0)
    pos, types, cell = random_system(rng, n_atoms, "periodic", side, types=(1, 6, 7, 8, 9))
    systems = [orc.System(pos, types, cell)]
    h = hypers(rc=6.0, lmax=lmax, nmax=nmax)
    oracle = oracle_to_tensormap(orc.soap_power_spectrum(systems, h, ["positions"]))
    L, n = lmax + 1, nmax + 1
    worst = {}
    for fast in (False, True):
        calc = ra.SoapPowerSpectrum(**h)
        calc.set_precision(32)
        calc.set_fast_form(fast)
        tm = calc.compute(ra_systems(systems), gradients=["positions"])
        compare_tensormaps(tm, oracle, 2e-4, 2e-4, gradients=["positions"])
        ev = per_l_pair_errors(tm, oracle, L, n)
        eg = per_l_pair_errors(tm, oracle, L, n, "positions")
        worst[fast] = (max(ev.values()), max(eg.values()))
        say("  fast" if fast else "  plain", "worst (l, type pair) error: values %.2e at %s, gradients %.2e at %s"
            % (worst[fast][0], max(ev, key=ev.get), worst[fast][1], max(eg, key=eg.get)))
        assert len(ev) == L * len(oracle.keys.values)
    assert max(worst[True]) <= 1e-4, worst[True]
    assert max(worst[False]) <= 1e-4, worst[False]


# ---------------------------------------------------------------------------- radial spectrum
RADIAL_SPECTRUM = {
    "cutoff": {"radius": 4.5, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"radial": {"type": "Gto", "max_radial": 7}, "spline_accuracy": 1e-8},
}


@pytest.mark.parametrize("precision", [64, 32])
def test_radial_spectrum_vs_oracle(precision):
    """SoapRadialSpectrum (the l = 0 expansion, relabelled) against the oracle: keys
    (center_type, neighbor_type), no components, values and position / strain / cell
    gradients; FP64 within 1e-10, FP32 within 2e-4 (the comparator of this file)."""
    rng = np.random.default_rng(61)
    systems = []
    for kind, n, L in [("periodic", 20, 6.5), ("triclinic", 16, 6.0), ("cluster", 12, 8.0)]:
        pos, types, cell = random_system(rng, n, kind, L, types=(1, 6, 8))
        systems.append(orc.System(pos, types, cell))
    grads = ["positions", "strain", "cell"]
    expected = oracle_to_tensormap(orc.soap_radial_spectrum(systems, RADIAL_SPECTRUM, grads))
    calc = ra.SoapRadialSpectrum(**RADIAL_SPECTRUM)
    calc.set_precision(precision)
    got = calc.compute(ra_systems(systems), gradients=grads)
    assert got.keys.names == ["center_type", "neighbor_type"]
    for block in got.blocks():
        assert len(block.components) == 0
        assert block.properties.names == ["n"]
    tol = 1e-10 if precision == 64 else 2e-4
    compare_tensormaps(got, expected, tol, tol, gradients=grads)
