"""GPU parity of the LODE spherical expansion against the reference's fixtures
(featomic/tests/data, generated for potential exponents 1 to 6) and against the FP64
oracle on random periodic systems.

Tolerances: fixtures as in the oracle's own fixture tests (relative 1e-8 with an
absolute floor of 1e-9 of the largest element, the reference uses 1e-6); FP64 kernels
against the oracle within |a - b| <= 1e-10 (|b| + max|b|); FP32 kernels (phases and
per-type sums still in FP64) within 1e-4 (|b| + max|b|).
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import (assert_relative_eq, compare_tensormaps, golden, load_npy_gz, oracle_to_tensormap, random_system,
                     sum_gradients)

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def ra_systems(systems):
    return [ra.System(s.types, s.positions, s.cell) for s in systems]


def flatten(tm):
    tm = tm.keys_to_samples("center_type").keys_to_properties("neighbor_type")
    return tm.components_to_properties("o3_mu").keys_to_properties("o3_lambda")


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_lode_fixture_values(p):
    say("LODE fixture values, exponent", p)
    systems, h = orc.load_input(golden(f"lode-spherical-expansion/exponent-{p}/values-input.json"))
    tm = ra.LodeSphericalExpansion(**h).compute(ra_systems(systems))
    expected = load_npy_gz(f"lode-spherical-expansion/exponent-{p}/values.npy.gz")
    assert_relative_eq(flatten(tm).block().values, expected, 1e-8, 1e-9 * np.abs(expected).max())
    compare_tensormaps(tm, oracle_to_tensormap(orc.lode_spherical_expansion(systems, h)), 1e-10, 1e-10)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_lode_fixture_gradients(p):
    say("LODE fixture gradients, exponent", p)
    systems, h = orc.load_input(golden(f"lode-spherical-expansion/exponent-{p}/gradients-input.json"))
    n_atoms = sum(len(s) for s in systems)
    tm = ra.LodeSphericalExpansion(**h).compute(ra_systems(systems), gradients=["positions"])
    expected = load_npy_gz(f"lode-spherical-expansion/exponent-{p}/positions-gradient.npy.gz")
    got = sum_gradients(n_atoms, flatten(tm).block().gradient("positions"))
    assert_relative_eq(got, expected, 1e-8, 1e-9 * np.abs(expected).max())


def lode_hypers(p, smearing=1.0, lmax=4, nmax=5, radius=3.0, accuracy=1e-8, weight=None):
    density = {"type": "SmearedPowerLaw", "smearing": smearing, "exponent": p}
    if weight is not None:
        density["center_atom_weight"] = weight
    return {"density": density,
            "basis": {"type": "TensorProduct", "max_angular": lmax,
                      "radial": {"type": "Gto", "max_radial": nmax, "radius": radius}, "spline_accuracy": accuracy}}


RANDOM_CASES = [
    # (seed, systems as (n atoms, kind, box), hypers)
    (0, [(12, "periodic", 6.0), (9, "triclinic", 5.5)], lode_hypers(1)),
    (1, [(10, "triclinic", 6.5)], lode_hypers(0, smearing=0.8, weight=0.5)),
    (2, [(8, "periodic", 5.0), (14, "periodic", 6.0)], lode_hypers(3, lmax=3, nmax=4)),
    (3, [(11, "triclinic", 6.0)], lode_hypers(4, smearing=1.2, accuracy=None)),
    (4, [(7, "periodic", 5.5)], lode_hypers(6, lmax=2, nmax=3, weight=0.0)),
]


def make_systems(seed, specs):
    rng = np.random.default_rng(1000 + seed)
    out = []
    for n, kind, box in specs:
        pos, types, cell = random_system(rng, n, kind, box, types=(1, 6, 8))
        out.append(orc.System(pos, types, cell))
    return out


@pytest.mark.parametrize("case", range(len(RANDOM_CASES)))
def test_lode_vs_oracle_fp64(case):
    seed, specs, h = RANDOM_CASES[case]
    say("LODE random case", case, specs, h["density"])
    systems = make_systems(seed, specs)
    calc = ra.LodeSphericalExpansion(**h)
    tm = calc.compute(ra_systems(systems), gradients=["positions"])
    oracle = oracle_to_tensormap(orc.lode_spherical_expansion(systems, h, ["positions"]))
    compare_tensormaps(tm, oracle, 1e-10, 1e-10, gradients=["positions"])


@pytest.mark.parametrize("case", [0, 3])
def test_lode_vs_oracle_fp32(case):
    seed, specs, h = RANDOM_CASES[case]
    say("LODE random case FP32", case)
    systems = make_systems(seed, specs)
    calc = ra.LodeSphericalExpansion(**h)
    calc.set_precision(32)
    tm = calc.compute(ra_systems(systems), gradients=["positions"])
    oracle = oracle_to_tensormap(orc.lode_spherical_expansion(systems, h, ["positions"]))
    compare_tensormaps(tm, oracle, 1e-4, 1e-4, gradients=["positions"])


def test_lode_errors():
    h = lode_hypers(1)
    calc = ra.LodeSphericalExpansion(**h)
    cluster = ra.System(np.array([1, 6]), np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), np.zeros((3, 3)))
    with pytest.raises(ra.FeatomicError, match="LODE can only be used with periodic systems"):
        calc.compute([cluster])
    tiny = lode_hypers(1, smearing=1.0)
    tiny["k_cutoff"] = 0.1
    calc = ra.LodeSphericalExpansion(**tiny)
    system = ra.System(np.array([1, 6]), np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), 5.0 * np.eye(3))
    with pytest.raises(ra.FeatomicError, match="No k-vectors"):
        calc.compute([system])
