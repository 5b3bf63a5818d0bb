"""The FP64 oracle against every golden fixture the reference's own tests hold.

Tolerances are the reference's (featomic/tests/*.rs) and, where the fixtures'
10-significant-digit rounding allows, tighter.
"""

import json

import numpy as np
import pytest

import featomic_oracle as orc
from helpers import (assert_relative_eq, flatten_ps, flatten_se, golden, load_npy_gz,
                     oracle_to_tensormap, sum_gradients)


def test_spherical_harmonics_fixture():
    # featomic/tests/spherical-harmonics.rs:10-36, tolerance 1e-11
    with open(golden("spherical-harmonics-input.json")) as fd:
        data = json.load(fd)
    expected = load_npy_gz("spherical-harmonics-values.npy.gz")
    lmax = data["max_angular"]
    for i, d in enumerate(data["directions"]):
        d = np.array(d) / np.linalg.norm(d)
        values, _ = orc.spherical_harmonics(lmax, d)
        for l in range(lmax + 1):
            assert_relative_eq(values[l * l:(l + 1) * (l + 1)], expected[i, l, :2 * l + 1], 1e-11, 1e-11)


def test_se_values_no_pbc():
    systems, hypers = orc.load_input(golden("spherical-expansion-values-input.json"))
    tm = flatten_se(oracle_to_tensormap(orc.soap_spherical_expansion(systems, hypers)))
    assert len(tm) == 1
    expected = load_npy_gz("spherical-expansion-values.npy.gz")
    # The reference's own tolerance for this fixture is 5e-5 (soap-spherical-expansion.rs:32):
    # the fixture predates the 30-term asymptotic 1F1 series (math/hyp1f1.rs:205-209), so
    # elements at large z carry ~1e-5 errors. 97% of elements agree to 1e-9.
    # (absolute floor: 1e-9 of the largest element)
    assert_relative_eq(tm.block().values, expected, 5e-5, 1e-9 * np.abs(expected).max())
    close = np.abs(tm.block().values - expected) <= 1e-9 * np.abs(expected) + 1e-12
    assert close.mean() > 0.95


def test_se_values_pbc():
    systems, hypers = orc.load_input(golden("spherical-expansion-pbc-values-input.json"))
    tm = flatten_se(oracle_to_tensormap(orc.soap_spherical_expansion(systems, hypers)))
    expected = load_npy_gz("spherical-expansion-pbc-values.npy.gz")
    assert_relative_eq(tm.block().values, expected, 1e-9, 1e-12)


def test_se_gradients():
    systems, hypers = orc.load_input(golden("spherical-expansion-gradients-input.json"))
    n_atoms = sum(len(s) for s in systems)
    tm = flatten_se(oracle_to_tensormap(orc.soap_spherical_expansion(systems, hypers, ["positions", "strain", "cell"])))
    block = tm.block()
    assert_relative_eq(sum_gradients(n_atoms, block.gradient("positions")),
                       load_npy_gz("spherical-expansion-positions-gradient.npy.gz"), 1e-8, 1e-12)
    assert_relative_eq(block.gradient("strain").values, load_npy_gz("spherical-expansion-strain-gradient.npy.gz"), 1e-8, 1e-12)
    assert_relative_eq(block.gradient("cell").values, load_npy_gz("spherical-expansion-cell-gradient.npy.gz"), 1e-8, 1e-12)


def test_ps_values():
    systems, hypers = orc.load_input(golden("soap-power-spectrum-values-input.json"))
    tm = flatten_ps(oracle_to_tensormap(orc.soap_power_spectrum(systems, hypers)))
    expected = load_npy_gz("soap-power-spectrum-values.npy.gz")
    # Same hyper-parameters as the SE no-pbc fixture (large-z 1F1 history); reference
    # tolerance is 1e-5 (soap-power-spectrum.rs:28). >99% of elements agree to 1e-9.
    assert_relative_eq(tm.block().values, expected, 1e-5, 1e-12)
    close = np.abs(tm.block().values - expected) <= 1e-9 * np.abs(expected) + 1e-12
    assert close.mean() > 0.99


def test_ps_gradients():
    systems, hypers = orc.load_input(golden("soap-power-spectrum-gradients-input.json"))
    n_atoms = sum(len(s) for s in systems)
    tm = flatten_ps(oracle_to_tensormap(orc.soap_power_spectrum(systems, hypers, ["positions", "strain", "cell"])))
    block = tm.block()
    assert_relative_eq(sum_gradients(n_atoms, block.gradient("positions")),
                       load_npy_gz("soap-power-spectrum-positions-gradient.npy.gz"), 1e-8, 1e-12)
    assert_relative_eq(block.gradient("strain").values, load_npy_gz("soap-power-spectrum-strain-gradient.npy.gz"), 1e-8, 1e-12)
    assert_relative_eq(block.gradient("cell").values, load_npy_gz("soap-power-spectrum-cell-gradient.npy.gz"), 1e-8, 1e-12)


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_lode_values(p):
    systems, hypers = orc.load_input(golden(f"lode-spherical-expansion/exponent-{p}/values-input.json"))
    tm = oracle_to_tensormap(orc.lode_spherical_expansion(systems, hypers))
    tm = tm.keys_to_samples("center_type").keys_to_properties("neighbor_type")
    tm = tm.components_to_properties("o3_mu").keys_to_properties("o3_lambda")
    expected = load_npy_gz(f"lode-spherical-expansion/exponent-{p}/values.npy.gz")
    # featomic/tests/lode-spherical-expansion.rs:42 uses max_relative 1e-6; the k-space sums
    # leave cancellation noise of ~1e-10 of the block scale on elements that vanish by
    # symmetry, so the absolute floor is stated relative to the largest element.
    assert_relative_eq(tm.block().values, expected, 1e-8, 1e-9 * np.abs(expected).max())


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6])
def test_lode_gradients(p):
    systems, hypers = orc.load_input(golden(f"lode-spherical-expansion/exponent-{p}/gradients-input.json"))
    n_atoms = sum(len(s) for s in systems)
    tm = oracle_to_tensormap(orc.lode_spherical_expansion(systems, hypers, ["positions"]))
    tm = tm.keys_to_samples("center_type").keys_to_properties("neighbor_type")
    tm = tm.components_to_properties("o3_mu").keys_to_properties("o3_lambda")
    expected = load_npy_gz(f"lode-spherical-expansion/exponent-{p}/positions-gradient.npy.gz")
    assert_relative_eq(sum_gradients(n_atoms, tm.block().gradient("positions")), expected, 1e-8,
                       1e-9 * np.abs(expected).max())


def test_oracle_radial_spectrum_is_the_l0_expansion():
    """The radial spectrum (soap/radial_spectrum.rs:75-135) is the l = 0 block of the
    spherical expansion: the oracle's relabelled expansion with max_angular = 0 equals
    the l = 0 blocks of an expansion with max_angular = 3 (the l = 0 radial functions
    and spline do not depend on max_angular), values and all gradients."""
    import featomic_oracle as orc

    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 6.0, size=(10, 3))
    types = np.array([1, 6, 8, 1, 1, 6, 8, 1, 6, 1], dtype=np.int32)
    system = orc.System(pos, types, np.eye(3) * 6.0)
    radial = {"type": "Gto", "max_radial": 4}
    cutoff = {"radius": 4.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}}
    density = {"type": "Gaussian", "width": 0.4}
    grads = ["positions", "strain", "cell"]
    rs = orc.soap_radial_spectrum([system], {"cutoff": cutoff, "density": density,
                                             "basis": {"radial": radial}}, grads)
    se = orc.soap_spherical_expansion([system], {"cutoff": cutoff, "density": density,
                                                 "basis": {"type": "TensorProduct", "max_angular": 3, "radial": radial}},
                                      grads)
    l0 = [k for k, key in enumerate(se["keys"]) if key[0] == 0]
    assert [tuple(k) for k in rs["keys"]] == [tuple(se["keys"][k][2:]) for k in l0]
    for blk, k in zip(rs["blocks"], l0):
        ref = se["blocks"][k]
        assert np.array_equal(blk["samples"], ref["samples"])
        assert np.allclose(blk["values"], ref["values"][:, 0, :], rtol=1e-12, atol=1e-14)
        for param in grads:
            g, gr = blk["gradients"][param], ref["gradients"][param]
            assert np.array_equal(g["samples"], gr["samples"])
            assert np.allclose(g["values"], gr["values"][..., 0, :], rtol=1e-12, atol=1e-14)
