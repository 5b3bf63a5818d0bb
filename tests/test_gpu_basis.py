"""GPU tests of the Explicit basis, the Tabulated radial basis and the DiracDelta
density (featomic/src/calculators/shared/basis/mod.rs:9-116, basis/radial/tabulated.rs,
soap/radial_integral/mod.rs:78-183).

The kernels compute the tensor-product shape (largest channel, largest radial size)
with each channel's own spline and zero functions elsewhere; the host TensorMap
keeps the basis' own channels and properties. Checked against the FP64 oracle
(1e-10 in FP64, the element-wise bar in FP32) and, for tabulated channels made of the
GTO spline's knots, against the Gaussian + GTO calculator.
"""

import copy
import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import assert_block_approx, compare_tensormaps, oracle_to_tensormap, random_system

pytestmark = pytest.mark.gpu

CUTOFF = 4.5
GAUSS = {
    "cutoff": {"radius": CUTOFF, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5}},
}
EXPLICIT_SIZES = {0: 6, 2: 4, 3: 8}
CALCULATORS = {"spherical expansion": (ra.SphericalExpansion, orc.soap_spherical_expansion),
               "power spectrum": (ra.SoapPowerSpectrum, orc.soap_power_spectrum)}


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def systems(seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for kind, n in (("cluster", 18), ("periodic", 22), ("triclinic", 16)):
        pos, types, cell = random_system(rng, n, kind, 7.0, types=(1, 6, 8))
        out.append(orc.System(pos, types, cell))
    return out


def ra_systems(syst):
    return [ra.System(s.types, s.positions, s.cell) for s in syst]


def tabulated_from_gto(l, n_radial, max_angular):
    radial = orc.Radial.soap(CUTOFF, 0.3, max_angular, n_radial, 1e-8)
    pos, val, der = radial.knots(l)
    return {"type": "Tabulated", "points": [{"position": float(p), "values": v.tolist(), "derivatives": d.tolist()}
                                            for p, v, d in zip(pos, val, der)]}


def random_tabulated(seed=11, n=4):
    rng = np.random.default_rng(seed)
    positions = np.concatenate([[0.0], np.sort(rng.uniform(0.05, CUTOFF - 0.05, 40)), [CUTOFF]])
    # smooth-ish functions vanishing at the cutoff, with their derivatives
    values = np.stack([np.cos(positions * (k + 1) * 0.7) * (CUTOFF - positions) for k in range(n)], axis=1)
    derivs = np.stack([-(k + 1) * 0.7 * np.sin(positions * (k + 1) * 0.7) * (CUTOFF - positions)
                       - np.cos(positions * (k + 1) * 0.7) for k in range(n)], axis=1)
    return {"type": "Tabulated", "points": [{"position": float(p), "values": v.tolist(), "derivatives": d.tolist()}
                                            for p, v, d in zip(positions, values, derivs)]}


def hypers_with(density, basis):
    h = copy.deepcopy(GAUSS)
    h["density"] = density
    h["basis"] = basis
    return h


def explicit_gto():
    return hypers_with(GAUSS["density"], {"type": "Explicit", "by_angular": {
        str(l): {"type": "Gto", "max_radial": s - 1} for l, s in EXPLICIT_SIZES.items()}})


def compute(cls, hypers, syst, precision, gradients=("positions",)):
    calc = cls(**hypers)
    calc.set_precision(precision)
    return calc.compute(ra_systems(syst), gradients=list(gradients))


def check_fp32(tm, ref):
    assert np.array_equal(tm.keys.values, ref.keys.values)
    for bg, be in zip(tm.blocks(), ref.blocks()):
        assert bg.samples == be.samples
        assert_block_approx(bg.values, be.values, 1e-4, 1e-5, "values")
        assert_block_approx(bg.gradient("positions").values, be.gradient("positions").values, 1e-4, 3e-5, "gradients")


@pytest.mark.parametrize("kind", list(CALCULATORS))
@pytest.mark.parametrize("precision", [64, 32])
def test_tabulated_gto_knots_equal_gaussian_gto(kind, precision):
    """Every channel tabulated from the GTO spline's knots (the oracle's, which agree
    with the library's to the last bits of the special functions), with a DiracDelta
    density: the Gaussian + GTO result (featomic's SoapSpliner construction)."""
    cls, _ = CALCULATORS[kind]
    lmax, n = 4, 6
    by_angular = {str(l): tabulated_from_gto(l, n, lmax) for l in range(lmax + 1)}
    tab = hypers_with({"type": "DiracDelta"}, {"type": "Explicit", "by_angular": by_angular})
    syst = systems()
    say(kind, precision)
    ref = compute(cls, GAUSS, syst, precision)
    got = compute(cls, tab, syst, precision)
    if precision == 64:
        compare_tensormaps(got, ref, 1e-12, 1e-12, gradients=["positions"])
    else:
        compare_tensormaps(got, ref, 1e-5, 1e-6, gradients=["positions"])


@pytest.mark.parametrize("kind", list(CALCULATORS))
def test_tabulated_vs_oracle(kind):
    """A TensorProduct basis with one tabulated radial integral on 42 non-uniform points."""
    cls, oracle = CALCULATORS[kind]
    hypers = hypers_with({"type": "DiracDelta", "center_atom_weight": 0.7},
                         {"type": "TensorProduct", "max_angular": 3, "radial": random_tabulated()})
    syst = systems(5)
    ref = oracle_to_tensormap(oracle(syst, hypers, ["positions"]))
    say(kind, "FP64")
    compare_tensormaps(compute(cls, hypers, syst, 64), ref, 1e-10, 1e-10, gradients=["positions"])
    say(kind, "FP32")
    check_fp32(compute(cls, hypers, syst, 32), ref)


@pytest.mark.parametrize("kind", list(CALCULATORS))
def test_explicit_gto_vs_oracle(kind):
    """Explicit {0: 6, 2: 4, 3: 8 radial functions}: channels 1 absent, every channel its
    own GTO basis; keys and properties as the reference lists them."""
    cls, oracle = CALCULATORS[kind]
    hypers = explicit_gto()
    syst = systems(7)
    ref = oracle_to_tensormap(oracle(syst, hypers, ["positions"]))
    say(kind, "FP64")
    tm = compute(cls, hypers, syst, 64)
    compare_tensormaps(tm, ref, 1e-10, 1e-10, gradients=["positions"])
    if kind == "spherical expansion":
        assert sorted({int(k[0]) for k in tm.keys.values}) == sorted(EXPLICIT_SIZES)
        for key, block in zip(tm.keys.values, tm.blocks()):
            assert block.values.shape[-1] == EXPLICIT_SIZES[int(key[0])]
    else:
        props = [tuple(p) for p in list(tm.blocks())[0].properties.values]
        assert props == [(l, a, b) for l, s in sorted(EXPLICIT_SIZES.items()) for a in range(s) for b in range(s)]
    say(kind, "FP32")
    check_fp32(compute(cls, hypers, syst, 32), ref)


def test_explicit_by_pair_vs_oracle():
    hypers = explicit_gto()
    syst = systems(9)[:2]
    ref = oracle_to_tensormap(orc.soap_spherical_expansion_by_pair(syst, hypers, ["positions"]))
    calc = ra.SphericalExpansionByPair(**hypers)
    calc.set_precision(64)
    tm = calc.compute(ra_systems(syst), gradients=["positions"])
    compare_tensormaps(tm, ref, 1e-10, 1e-10, gradients=["positions"])


def test_explicit_property_selection():
    """Selected properties refer to the Explicit basis' own labels."""
    hypers = explicit_gto()
    syst = systems(7)
    calc = ra.SoapPowerSpectrum(**hypers)
    calc.set_precision(64)
    full = calc.compute(ra_systems(syst), gradients=["positions"])
    wanted = ra.Labels(["l", "n_1", "n_2"], np.array([[3, 7, 1], [0, 5, 5], [2, 3, 0]], dtype=np.int32))
    part = calc.compute(ra_systems(syst), gradients=["positions"], selected_properties=wanted)
    for bf, bp in zip(full.blocks(), part.blocks()):
        props = [tuple(p) for p in bf.properties.values]
        cols = [props.index(tuple(w)) for w in wanted.values]
        assert np.array_equal(bp.values, bf.values[:, cols])
        assert np.array_equal(bp.gradient("positions").values, bf.gradient("positions").values[..., cols])


def test_padded_basis_on_the_device_path_is_refused():
    from rascaline_amd.device import DeviceSystems, compute_device

    calc = ra.SoapPowerSpectrum(**explicit_gto())
    syst = systems(7)
    with pytest.raises(ra.FeatomicError, match="featomic_calculator_compute"):
        compute_device(calc, DeviceSystems([(s.positions, s.types, s.cell) for s in syst]), ["positions"])
    with pytest.raises(ra.FeatomicError, match="radial surrogate"):
        calc.train_radial_surrogate()
