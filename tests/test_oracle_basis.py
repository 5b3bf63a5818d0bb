"""Pins of the oracle's Explicit basis, Tabulated radial basis and DiracDelta density
(featomic/src/calculators/shared/basis/mod.rs:9-116, basis/radial/tabulated.rs,
soap/radial_integral/mod.rs:78-183).

* A tabulated radial integral is the Hermite cubic spline through its points: the
  oracle's values and derivatives match scipy's CubicHermiteSpline, an independent
  implementation, on a two-atom cluster (c_nlm = f_c(r) R_n(r) Y_lm(r^) for the
  neighbour).
* A DiracDelta density with, for every channel l, the GTO spline knots of channel l
  as its Tabulated points is the Gaussian + GTO expansion itself: the reference's
  own way of splining a radial integral (featomic's Python SoapSpliner), which must
  give the same numbers since it evaluates the same spline.
* An Explicit basis with GTO channels of different sizes gives, for each channel l
  with n_l functions, the block l (and the (l, n_1, n_2) power spectrum columns) of
  the TensorProduct basis with max_radial = n_l - 1.
"""

import copy

import numpy as np
import pytest
from scipy.interpolate import CubicHermiteSpline

import featomic_oracle as orc
from helpers import random_system

CUTOFF = 4.5
GAUSS = {
    "cutoff": {"radius": CUTOFF, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5}},
}


def tabulated_from_gto(l, n_radial, max_angular=None, accuracy=1e-8, width=0.3):
    """Tabulated points of channel l of the GTO x Gaussian spline (its knots)."""
    radial = orc.Radial.soap(CUTOFF, width, l if max_angular is None else max_angular, n_radial, accuracy)
    pos, val, der = radial.knots(l)
    return {"type": "Tabulated", "points": [{"position": float(p), "values": v.tolist(), "derivatives": d.tolist()}
                                            for p, v, d in zip(pos, val, der)]}


def dirac(hypers, basis):
    h = copy.deepcopy(hypers)
    h["density"] = {"type": "DiracDelta"}
    h["basis"] = basis
    return h


def systems(seed=3):
    rng = np.random.default_rng(seed)
    out = []
    for kind in ("cluster", "periodic"):
        pos, types, cell = random_system(rng, 14, kind, 7.0, types=(1, 8))
        out.append(orc.System(pos, types, cell))
    return out


def blocks_by_key(result):
    return {tuple(int(x) for x in k): b for k, b in zip(result["keys"], result["blocks"])}


def assert_blocks_equal(a, b, gradients=True):
    assert np.array_equal(a["samples"], b["samples"])
    assert np.array_equal(a["properties"], b["properties"])
    assert np.array_equal(a["values"], b["values"])
    if gradients:
        ga, gb = a["gradients"]["positions"], b["gradients"]["positions"]
        assert np.array_equal(ga["samples"], gb["samples"])
        assert np.array_equal(ga["values"], gb["values"])


def test_tabulated_spline_is_scipy_hermite():
    """Two-atom cluster: the neighbour's contribution to c_nlm is f_c(r) R_n(r) Y_lm(r^)
    with R the Hermite spline through the points (scipy), on a deliberately
    non-uniform, unsorted point list."""
    rng = np.random.default_rng(11)
    n = 3
    positions = np.concatenate([[0.0], np.sort(rng.uniform(0.1, CUTOFF - 0.1, 9)), [CUTOFF]])
    values = rng.normal(size=(len(positions), n))
    derivatives = rng.normal(size=(len(positions), n))
    order = rng.permutation(len(positions))  # the reference sorts the points
    radial = {"type": "Tabulated", "points": [{"position": float(positions[k]), "values": values[k].tolist(),
                                                "derivatives": derivatives[k].tolist()} for k in order]}
    lmax = 3
    hypers = dirac(GAUSS, {"type": "TensorProduct", "max_angular": lmax, "radial": radial})
    hypers["cutoff"]["smoothing"] = {"type": "Step"}
    spline = CubicHermiteSpline(positions, values, derivatives)
    for r in (0.37, 1.9, 3.3):
        direction = np.array([0.3, -0.5, 0.8])
        direction /= np.linalg.norm(direction)
        pos = np.array([[0.0, 0.0, 0.0], r * direction])
        system = orc.System(pos, np.array([1, 8], dtype=np.int32), np.zeros((3, 3)))
        se = orc.soap_spherical_expansion([system], hypers)
        y, _ = orc.spherical_harmonics(lmax, direction, gradients=False)
        for key, block in blocks_by_key(se).items():
            l, _, ct, nt = key
            if ct != 1 or nt != 8:
                continue
            row = list(map(tuple, block["samples"])).index((0, 0))
            expected = np.outer(y[l * l:(l + 1) * (l + 1)], spline(r))
            assert np.allclose(block["values"][row], expected, rtol=1e-12, atol=1e-14), (r, l)


def test_dirac_tabulated_reproduces_gaussian_gto():
    """Explicit basis, channel l = the GTO knots of channel l, DiracDelta density: the
    same spherical expansion and power spectrum, bit for bit."""
    syst = systems()
    lmax, n = GAUSS["basis"]["max_angular"], GAUSS["basis"]["radial"]["max_radial"] + 1
    by_angular = {str(l): tabulated_from_gto(l, n, max_angular=lmax) for l in range(lmax + 1)}
    hypers = dirac(GAUSS, {"type": "Explicit", "by_angular": by_angular})
    for fn in (orc.soap_spherical_expansion, orc.soap_power_spectrum):
        ref = blocks_by_key(fn(syst, GAUSS, ["positions"]))
        got = blocks_by_key(fn(syst, hypers, ["positions"]))
        assert ref.keys() == got.keys()
        for key in ref:
            assert_blocks_equal(got[key], ref[key])


def test_tensor_product_tabulated_channel_zero():
    """TensorProduct with one tabulated radial basis (the GTO knots of l = 0) uses it
    for every l: the l = 0 blocks equal the Gaussian + GTO ones."""
    syst = systems(5)
    n = GAUSS["basis"]["radial"]["max_radial"] + 1
    hypers = dirac(GAUSS, {"type": "TensorProduct", "max_angular": 2, "radial": tabulated_from_gto(0, n, max_angular=4)})
    ref = blocks_by_key(orc.soap_spherical_expansion(syst, GAUSS, ["positions"]))
    got = blocks_by_key(orc.soap_spherical_expansion(syst, hypers, ["positions"]))
    assert {k for k in got} == {k for k in ref if k[0] <= 2}
    for key in got:
        if key[0] == 0:
            assert_blocks_equal(got[key], ref[key])


def test_explicit_gto_channels():
    """Explicit {0: 6 functions, 2: 4, 3: 8}: block l equals the TensorProduct block l
    with max_radial = n_l - 1 (spherical expansion), and the (l, n_1, n_2) columns of
    its power spectrum (power_spectrum.rs:498-523)."""
    syst = systems(7)
    sizes = {0: 6, 2: 4, 3: 8}
    explicit = copy.deepcopy(GAUSS)
    explicit["basis"] = {"type": "Explicit",
                         "by_angular": {str(l): {"type": "Gto", "max_radial": s - 1} for l, s in sizes.items()}}
    se = blocks_by_key(orc.soap_spherical_expansion(syst, explicit, ["positions"]))
    ps = blocks_by_key(orc.soap_power_spectrum(syst, explicit, ["positions"]))
    assert {k[0] for k in se} == set(sizes)
    for l, s in sizes.items():
        tp = copy.deepcopy(GAUSS)
        tp["basis"] = {"type": "TensorProduct", "max_angular": 3, "radial": {"type": "Gto", "max_radial": s - 1}}
        se_tp = blocks_by_key(orc.soap_spherical_expansion(syst, tp, ["positions"]))
        for key, block in se.items():
            if key[0] == l:
                assert_blocks_equal(block, se_tp[key])
        ps_tp = blocks_by_key(orc.soap_power_spectrum(syst, tp, ["positions"]))
        for key, block in ps.items():
            props = [tuple(p) for p in block["properties"]]
            cols = [k for k, p in enumerate(props) if p[0] == l]
            tp_props = [tuple(p) for p in ps_tp[key]["properties"]]
            tp_cols = [tp_props.index(props[k]) for k in cols]
            assert len(cols) == s * s
            assert np.array_equal(block["values"][:, cols], ps_tp[key]["values"][:, tp_cols])
            g, gt = block["gradients"]["positions"], ps_tp[key]["gradients"]["positions"]
            assert np.array_equal(g["values"][..., cols], gt["values"][..., tp_cols])
    # properties of the power spectrum: channel by channel, n_1 and n_2 below the channel's size
    props = [tuple(p) for p in next(iter(ps.values()))["properties"]]
    assert props == [(l, a, b) for l, s in sorted(sizes.items()) for a in range(s) for b in range(s)]
