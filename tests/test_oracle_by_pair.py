"""Oracle SphericalExpansionByPair (soap/spherical_expansion_pair.rs) pinned by the
reference's own checks of that calculator:

* ``sums_to_spherical_expansion`` (spherical_expansion_pair.rs:1000-1040): summing the
  pairs over the second atom gives the spherical expansion, whose oracle is pinned by
  the reference's fixtures; here for values and for position, strain and cell
  gradients, on the reference's water + methane systems and a random batch;
* ``finite_differences_positions / cell / strain`` (spherical_expansion_pair.rs:
  889-933) on the water system with the reference's parameters, displacement 1e-6,
  max_relative 1e-5.
"""

import numpy as np
import pytest

import featomic_oracle as orc
from helpers import random_system

# spherical_expansion_pair.rs:861-887
PARAMS = {
    "cutoff": {"radius": 7.3, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3,
                "scaling": {"type": "Willatt2018", "scale": 1.5, "rate": 0.8, "exponent": 2.0}},
    "basis": {"type": "TensorProduct", "max_angular": 2, "radial": {"type": "Gto", "max_radial": 5},
              "spline_accuracy": 1e-8},
}


def water():
    # featomic/src/systems/test_utils.rs:29-36
    return orc.System([[0.0, 0.0, 0.0], [0.0, 0.75545, -0.58895], [0.0, -0.75545, -0.58895]], [-42, 1, 1],
                      np.eye(3) * 3.0)


def methane():
    # test_utils.rs:19-27
    return orc.System([[5.0, 5.0, 5.0], [5.5288, 5.1610, 5.9359], [5.2051, 5.8240, 4.3214],
                       [5.3345, 4.0686, 4.5504], [3.9315, 4.9463, 5.1921]], [6, 1, 1, 1, 1], np.eye(3) * 5.0)


def by_key(tm):
    return {tuple(int(x) for x in k): b for k, b in zip(np.asarray(tm["keys"]), tm["blocks"])}


@pytest.mark.parametrize("case", ["reference", "random"])
def test_pairs_sum_to_spherical_expansion(case):
    if case == "reference":
        systems, hypers = [water(), methane()], PARAMS
    else:
        rng = np.random.default_rng(8)
        systems = [orc.System(*random_system(rng, 12, "triclinic", 6.0)),
                   orc.System(*random_system(rng, 9, "cluster", 7.0))]
        hypers = dict(PARAMS, cutoff={"radius": 4.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
                      basis={"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 4}})
    grads = ["positions", "strain", "cell"]
    pair = by_key(orc.soap_spherical_expansion_by_pair(systems, hypers, grads))
    se = by_key(orc.soap_spherical_expansion(systems, hypers, grads))
    # the same keys (spherical_expansion_pair.rs:1010-1016), other order
    assert sorted(pair) == sorted(se)
    n_atoms = [len(s) for s in systems]
    for key, blk in se.items():
        pb = pair[key]
        index = {(int(s), int(a)): k for k, (s, a) in enumerate(blk["samples"])}
        summed = np.zeros_like(blk["values"])
        for k, smp in enumerate(pb["samples"]):
            summed[index[(int(smp[0]), int(smp[1]))]] += pb["values"][k]
        scale = max(np.abs(blk["values"]).max(), 1e-300)
        assert np.abs(summed - blk["values"]).max() <= 1e-12 * scale, key
        # position gradients: dense [sample][system atom] maps
        g_se = {}
        for r, (s_i, system, atom) in enumerate(blk["gradients"]["positions"]["samples"]):
            g_se[(int(s_i), int(atom))] = blk["gradients"]["positions"]["values"][r]
        g_pair = {}
        pg = pb["gradients"]["positions"]
        for r, (k, system, atom) in enumerate(pg["samples"]):
            center = index[(int(system), int(pb["samples"][k][1]))]
            g_pair.setdefault((center, int(atom)), 0.0)
            g_pair[(center, int(atom))] = g_pair[(center, int(atom))] + pg["values"][r]
        gscale = max(max((np.abs(v).max() for v in g_se.values()), default=0.0), 1e-300)
        for where in set(g_se) | set(g_pair):
            a = g_pair.get(where, 0.0)
            b = g_se.get(where, 0.0)
            assert np.abs(np.asarray(a) - np.asarray(b)).max() <= 1e-12 * gscale, (key, where)
        for param in ("strain", "cell"):
            summed = np.zeros_like(blk["gradients"][param]["values"])
            for k, smp in enumerate(pb["samples"]):
                summed[index[(int(smp[0]), int(smp[1]))]] += pb["gradients"][param]["values"][k]
            ref = blk["gradients"][param]["values"]
            assert np.abs(summed - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300), (key, param)
    assert sum(n_atoms) > 0


def _assert_fd(fd, analytic, max_relative, epsilon):
    diff = np.abs(fd - analytic)
    bound = np.maximum(epsilon, max_relative * np.maximum(np.abs(fd), np.abs(analytic)))
    assert np.all(diff <= bound), float((diff / np.maximum(bound, 1e-300)).max())


def _values_by_sample(tm):
    return [{tuple(int(x) for x in s): b["values"][k] for k, s in enumerate(b["samples"])} for b in tm["blocks"]]


def test_finite_differences_positions():
    system = water()
    ref = orc.soap_spherical_expansion_by_pair([system], PARAMS, ["positions"])
    h = 1e-6
    for atom in range(len(system)):
        for x in range(3):
            moved = []
            for sign in (0.5, -0.5):
                pos = system.positions.copy()
                pos[atom, x] += sign * h
                moved.append(_values_by_sample(orc.soap_spherical_expansion_by_pair(
                    [orc.System(pos, system.types, system.cell)], PARAMS)))
            for b, blk in enumerate(ref["blocks"]):
                g = blk["gradients"]["positions"]
                for r, (k, _, a) in enumerate(g["samples"]):
                    if a != atom:
                        continue
                    smp = tuple(int(v) for v in blk["samples"][k])
                    fd = (moved[0][b][smp] - moved[1][b][smp]) / h
                    _assert_fd(fd, g["values"][r, x], 1e-5, 1e-16)


@pytest.mark.parametrize("param", ["cell", "strain"])
def test_finite_differences_cell_and_strain(param):
    system = water()
    ref = orc.soap_spherical_expansion_by_pair([system], PARAMS, [param])
    h = 1e-6
    for a in range(3):
        for b in range(3):
            moved = []
            for sign in (0.5, -0.5):
                if param == "cell":
                    cell = system.cell.copy()
                    cell[a, b] += sign * h
                    pos = system.positions
                else:
                    strain = np.eye(3)
                    strain[a, b] += sign * h
                    cell = system.cell @ strain
                    pos = system.positions @ strain
                moved.append(_values_by_sample(orc.soap_spherical_expansion_by_pair(
                    [orc.System(pos, system.types, cell)], PARAMS)))
            for bi, blk in enumerate(ref["blocks"]):
                g = blk["gradients"][param]["values"]
                for k, smp in enumerate(blk["samples"]):
                    smp = tuple(int(v) for v in smp)
                    fd = (moved[0][bi][smp] - moved[1][bi][smp]) / h
                    _assert_fd(fd, g[k, a, b], 1e-5, 1e-9)
