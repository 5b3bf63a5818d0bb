"""Pins of the oracle paths that the golden fixtures leave open.

* LODE central-atom term, density Fourier transform and k = 0 term against the
  pyLODE values the reference's unit tests hold
  (lode/radial_integral/gto.rs:266-287, lode/spherical_expansion.rs:987-1082),
  at the reference's tolerances;
* LODE with exponent 0 against SOAP with a step cutoff
  (featomic/tests/lode-vs-soap.rs:8-80), which pins the p = 0 density and its
  k = 0 term through a second, independent calculator;
* Madelung constants of four crystals (featomic/tests/lode-madelung.rs:76-173),
  computed with center_atom_weight = 0, which pins the central-atom term as it
  enters the expansion;
* the Willatt2018 radial scaling (shared/density.rs:107-129) by its closed
  form: on a two-atom cluster the scaled expansion is s(r) times the unscaled
  one, and the position gradient picks up s'(r) r_hat times the unscaled values.
"""

import numpy as np
import pytest

import featomic_oracle as orc
from helpers import assert_relative_eq


def lode_hypers(exponent, smearing, radius, max_radial, max_angular=0, spline_accuracy=1e-8, center_atom_weight=1.0,
                k_cutoff=None):
    h = {
        "density": {"type": "SmearedPowerLaw", "smearing": smearing, "exponent": exponent,
                    "center_atom_weight": center_atom_weight},
        "basis": {"type": "TensorProduct", "max_angular": max_angular,
                  "radial": {"type": "Gto", "max_radial": max_radial, "radius": radius},
                  "spline_accuracy": spline_accuracy},
    }
    if k_cutoff is not None:
        h["k_cutoff"] = k_cutoff
    return h


# lode/radial_integral/gto.rs:269-274 (values from pyLODE)
PYLODE_CENTER = {
    0: [7.09990773e-01, 6.13767550e-01, 3.34161655e-01, 8.35301652e-02, 1.78439072e-02, -3.44944648e-05],
    1: [1.69193719, 2.02389574, 2.85086136, 3.84013091, 1.62869125, 7.03338899],
    2: [1.00532822, 1.10024472, 1.34843326, 1.19816598, 0.69150744, 1.2765415],
    6: [0.03811939, 0.03741200, 0.03115835, 0.01364843, 0.00534184, 0.00205973],
}


@pytest.mark.parametrize("exponent", sorted(PYLODE_CENTER))
def test_lode_center_contribution_pylode(exponent):
    got = orc.lode_center_contribution(lode_hypers(exponent, 1.0, 5.0, 5))
    assert_relative_eq(got, np.array(PYLODE_CENTER[exponent]), 3e-6)


@pytest.mark.parametrize("exponent,expected", [(0, 6.67432573), (4, 7.87480497), (6, 0.65623375)])
def test_lode_density_fourier_pylode(exponent, expected):
    # lode/spherical_expansion.rs:987-1025: |k| = 1e-12 and 1e-11, smearing 1
    got = orc.lode_density_fourier(lode_hypers(exponent, 1.0, 3.5, 5, 5), [1e-12, 1e-11])
    assert_relative_eq(got, np.array([expected, expected]), 1e-8)


@pytest.mark.parametrize("exponent,expected", [
    (0, [0.49695, 0.78753, 1.07009, 3.13526, -0.18495, 8.9746]),
    (6, [0.13337, 0.21136, 0.28719, 0.84143, -0.04964, 2.40858]),
])
def test_lode_k0_contribution_pylode(exponent, expected):
    # lode/spherical_expansion.rs:1029-1082, smearing 0.8, radius 3.5, 6 radial functions
    got = orc.lode_k0_contribution(lode_hypers(exponent, 0.8, 3.5, 5, 5))
    assert_relative_eq(got, np.array(expected), 1e-4)


@pytest.mark.parametrize("exponent", [1, 2, 3])
def test_lode_k0_contribution_vanishes(exponent):
    # compute_k0_contributions is zero for 0 < p <= 3 (lode/spherical_expansion.rs:352-372)
    assert np.all(orc.lode_k0_contribution(lode_hypers(exponent, 0.8, 3.5, 5, 5)) == 0.0)


def lode_vs_soap_case():
    """featomic/tests/lode-vs-soap.rs:8-80 (release-build sizes)."""
    pos = np.array([[1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [2.0, 2.2, 1.0], [2.3, 2.0, 1.5]])
    system = orc.System(pos, np.array([8, 8, 8, 8], dtype=np.int32), np.eye(3) * 20.0)
    max_radial, max_angular = 5, 2
    lode = {
        "k_cutoff": 16.0,
        "density": {"type": "SmearedPowerLaw", "smearing": 0.3, "exponent": 0},
        "basis": {"type": "TensorProduct", "max_angular": max_angular,
                  "radial": {"max_radial": max_radial, "type": "Gto", "radius": 3.0}, "spline_accuracy": None},
    }
    soap = {
        "cutoff": {"radius": 3.0, "smoothing": {"type": "Step"}},
        "density": {"type": "Gaussian", "width": 0.3},
        "basis": {"type": "TensorProduct", "max_angular": max_angular,
                  "radial": {"max_radial": max_radial, "type": "Gto"}, "spline_accuracy": None},
    }
    return system, lode, soap


def test_lode_p0_equals_soap_step():
    system, lode, soap = lode_vs_soap_case()
    got = orc.lode_spherical_expansion([system], lode)
    expected = orc.soap_spherical_expansion([system], soap)
    assert np.array_equal(got["keys"], expected["keys"])
    for a, b in zip(got["blocks"], expected["blocks"]):
        assert_relative_eq(a["values"], b["values"], 1e-4)


# featomic/tests/lode-madelung.rs:24-75
def nacl():
    cell = np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    return orc.System(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]), np.array([11, 17], dtype=np.int32), cell)


def cscl():
    return orc.System(np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]), np.array([17, 55], dtype=np.int32), np.eye(3))


def zns():
    cell = np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    return orc.System(np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]), np.array([16, 30], dtype=np.int32), cell)


def znso4():
    u = 3.0 / 8.0
    c = np.sqrt(1.0 / u)
    cell = np.array([[0.5, -0.5 * np.sqrt(3.0), 0.0], [0.5, 0.5 * np.sqrt(3.0), 0.0], [0.0, 0.0, c]])
    pos = np.array([
        [0.5, 0.5 / np.sqrt(3.0), 0.0],
        [0.5, 0.5 / np.sqrt(3.0), u * c],
        [0.5, -0.5 / np.sqrt(3.0), 0.5 * c],
        [0.5, -0.5 / np.sqrt(3.0), (0.5 + u) * c],
    ])
    return orc.System(pos, np.array([16, 30, 16, 30], dtype=np.int32), cell)


CRYSTALS = [
    (nacl, 1.7476),
    (cscl, 2.0 * 1.7626 / np.sqrt(3.0)),
    (zns, 2.0 * 1.6381 / np.sqrt(3.0)),
    (znso4, 1.6413 / np.sqrt(3.0 / 8.0)),
]


def madelung(system, hypers, gto_radius):
    out = orc.lode_spherical_expansion([system], hypers)
    factor = -1.0 / (4.0 * np.pi * gto_radius**2) ** 0.75
    # blocks 0 and 1: (center_type, neighbor_type) = (t0, t0) and (t0, t1) at l = 0; charges +1, -1
    return factor * (out["blocks"][0]["values"][0, 0, 0] - out["blocks"][1]["values"][0, 0, 0])


@pytest.mark.parametrize("gto_radius", [0.01, 0.027, 0.074, 0.2])
@pytest.mark.parametrize("smearing", [0.2, 0.1])
def test_madelung(gto_radius, smearing):
    # lode-madelung.rs:76-123, max_relative 8e-2, center_atom_weight 0
    hypers = lode_hypers(1, smearing, gto_radius, 0, 0, 1e-8, 0.0)
    for build, constant in CRYSTALS:
        got = madelung(build(), hypers, gto_radius)
        assert abs(got - constant) <= 8e-2 * max(abs(got), abs(constant)), (build.__name__, got, constant)


def test_madelung_high_accuracy():
    # lode-madelung.rs:129-173, max_relative 5e-5
    hypers = lode_hypers(1, 0.1, 0.01, 0, 0, 1e-8, 0.0, k_cutoff=50.0)
    for build, constant in CRYSTALS:
        got = madelung(build(), hypers, 0.01)
        assert abs(got - constant) <= 5e-5 * max(abs(got), abs(constant)), (build.__name__, got, constant)


def test_madelung_needs_the_center_term():
    # with center_atom_weight = 1 the central-atom term is not removed and the constant is far off:
    # the pin above really exercises that term
    hypers = lode_hypers(1, 0.1, 0.01, 0, 0, 1e-8, 1.0, k_cutoff=50.0)
    got = madelung(nacl(), hypers, 0.01)
    assert abs(got - 1.7476) > 0.1


def willatt(r, rate, scale, exponent):
    return rate / (rate + (r / scale) ** exponent)


def willatt_gradient(r, rate, scale, exponent):
    # d/dr of rate / (rate + (r/scale)^exponent)
    rs = (r / scale) ** exponent
    return -rate * exponent * rs / r / (rate + rs) ** 2


def soap_hypers(scaling=None, center_atom_weight=1.0):
    h = {
        "cutoff": {"radius": 4.5, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
        "density": {"type": "Gaussian", "width": 0.4, "center_atom_weight": center_atom_weight},
        "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5}},
    }
    if scaling is not None:
        h["density"]["scaling"] = scaling
    return h


@pytest.mark.parametrize("rate,scale,exponent", [(1.0, 2.0, 4.0), (0.3, 1.5, 2.5), (2.0, 3.0, 7.0)])
def test_willatt2018_scaling_closed_form(rate, scale, exponent):
    # two atoms of different types: each center's block of the other type holds only the pair term
    r_vec = np.array([1.1, -0.7, 2.3])
    r = np.linalg.norm(r_vec)
    system = orc.System(np.array([[0.0, 0.0, 0.0], r_vec]), np.array([1, 8], dtype=np.int32), np.zeros((3, 3)))
    scaling = {"type": "Willatt2018", "rate": rate, "scale": scale, "exponent": exponent}
    plain = orc.soap_spherical_expansion([system], soap_hypers(), ["positions"])
    scaled = orc.soap_spherical_expansion([system], soap_hypers(scaling), ["positions"])
    s = willatt(r, rate, scale, exponent)
    ds = willatt_gradient(r, rate, scale, exponent)
    keys = [tuple(k) for k in plain["keys"]]
    checked = 0
    for b, key in enumerate(keys):
        _, _, ct, nt = key
        if ct == nt:
            continue
        vp, vs = plain["blocks"][b]["values"], scaled["blocks"][b]["values"]
        assert_relative_eq(vs, s * vp, 1e-12, 1e-15)
        gp = plain["blocks"][b]["gradients"]["positions"]
        gs = scaled["blocks"][b]["gradients"]["positions"]
        assert np.array_equal(gp["samples"], gs["samples"])
        for row, (sample, _, atom) in enumerate(gp["samples"]):
            center = plain["blocks"][b]["samples"][sample][1]
            # d/dr_atom of s(|r_other - r_center|) c: the distance moves with +r_hat for the neighbour
            sign = 1.0 if atom != center else -1.0
            direction = (r_vec if center == 0 else -r_vec) / r
            expected = s * gp["values"][row] + sign * ds * direction[:, None, None] * vp[sample][None]
            assert_relative_eq(gs["values"][row], expected, 1e-10, 1e-14)
            checked += 1
    assert checked == 2 * 5 * 2  # two cross-type blocks per l (l <= 4), two rows each


def test_willatt2018_large_rate_is_unscaled():
    system = orc.System(np.array([[0.0, 0.0, 0.0], [1.2, 0.4, -0.9], [0.3, 2.0, 0.5]]),
                        np.array([1, 8, 8], dtype=np.int32), np.zeros((3, 3)))
    scaling = {"type": "Willatt2018", "rate": 1e12, "scale": 1.0, "exponent": 3.0}
    plain = orc.soap_spherical_expansion([system], soap_hypers())
    scaled = orc.soap_spherical_expansion([system], soap_hypers(scaling))
    for a, b in zip(plain["blocks"], scaled["blocks"]):
        # s = 1 - O(r^3 / rate): 1e-11 at these distances, below FP64 cancellation in small elements
        assert_relative_eq(b["values"], a["values"], 1e-10, 1e-12 * np.abs(a["values"]).max())


def test_lode_center_subset_is_the_full_computation():
    # orc_lode_system_centers (gradient rows of chosen centers, used for large cells) gives the
    # same rows as the full computation
    rng = np.random.default_rng(5)
    from helpers import random_system

    pos, types, cell = random_system(rng, 9, "triclinic", 5.5)
    system = orc.System(pos, types, cell)
    hypers = lode_hypers(1, 1.0, 3.0, 3, 2)
    species = np.unique(types)
    values, grads, self_grads = orc.lode_dense(system, hypers, species, centers=range(len(types)))
    sub_values, sub_grads, sub_self = orc.lode_dense(system, hypers, species, centers=[4, 1])
    assert np.array_equal(values, sub_values)
    assert np.array_equal(grads[[4, 1]], sub_grads)
    assert np.array_equal(self_grads[[4, 1]], sub_self)
    full = orc.lode_spherical_expansion([system], hypers, ["positions"])
    # block (center type, neighbour type) l = 0 row of center 4 with neighbour j = 0 (if of that type)
    t4, t0 = types[4], types[0]
    b = [tuple(k) for k in full["keys"]].index((0, 1, t4, t0))
    g = full["blocks"][b]["gradients"]["positions"]
    samples = [tuple(s) for s in full["blocks"][b]["samples"]]
    s4 = samples.index((0, 4))
    row = [tuple(r) for r in g["samples"]].index((s4, 0, 0))
    assert np.allclose(g["values"][row][:, 0, :], sub_grads[0, 0, :, 0, :], rtol=0, atol=0)
