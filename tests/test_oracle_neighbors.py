"""Oracle neighbour list against the reference's unit tests and a brute force."""

import itertools
import math

import numpy as np
import pytest

import featomic_oracle as orc


def test_non_periodic_ase_distances():
    # featomic/src/systems/neighbors.rs:413-444 (reference computed with ASE)
    pos = np.array([[0.134, 1.282, 1.701], [-0.273, 1.026, -1.471], [1.922, -0.124, 1.900],
                    [1.400, -0.464, 0.480], [0.149, 1.865, 0.635]])
    ref = [(0, 1, 3.2082345612501593), (0, 2, 2.283282943482914), (0, 3, 2.4783286706972505),
           (0, 4, 1.215100818862369), (1, 3, 2.9707625283755013), (1, 4, 2.3059143522689647),
           (2, 3, 1.550639867925496), (2, 4, 2.9495550511899244), (3, 4, 2.6482573515427084)]
    pairs = orc.neighbors(pos, np.zeros((3, 3)), 3.42)
    assert len(pairs) == len(ref)
    for p, (a, b, d) in zip(pairs, ref):
        assert (p["first"], p["second"]) == (a, b)
        assert p["distance"] == pytest.approx(d, rel=4e-16)


def test_fcc_self_images():
    # neighbors.rs:446-473
    cell = np.array([[0.0, 1.5, 1.5], [1.5, 0.0, 1.5], [1.5, 1.5, 0.0]])
    pairs = orc.neighbors(np.zeros((1, 3)), cell, 3.0)
    expected = [([1.0, 0.0, -1.0], [-1, 0, 1]), ([1.0, -1.0, 0.0], [-1, 1, 0]), ([0.0, 1.0, -1.0], [0, -1, 1]),
                ([1.0, 1.0, 0.0], [0, 0, 1]), ([1.0, 0.0, 1.0], [0, 1, 0]), ([0.0, 1.0, 1.0], [1, 0, 0])]
    assert len(pairs) == 6
    got = sorted((tuple(p["shift"]), tuple(p["vector"] / 1.5)) for p in pairs)
    exp = sorted((tuple(s), tuple(v)) for v, s in expected)
    for (gs, gv), (es, ev) in zip(got, exp):
        assert gs == es
        np.testing.assert_allclose(gv, ev, atol=1e-15)
    for p in pairs:
        assert p["distance"] == pytest.approx(2.1213203435596424, rel=4e-16)


def test_large_cell_small_cutoff():
    # neighbors.rs:475-503
    pos = np.array([[0, 0, 0], [0, 2, 0], [0, 0, 2], [-6, 0, 0], [-6, -2, 0], [-6, 0, -2.0]])
    pairs = orc.neighbors(pos, np.eye(3) * 54.0, 2.1)
    assert [(p["first"], p["second"]) for p in pairs] == [(0, 1), (0, 2), (3, 4), (3, 5)]
    for p in pairs:
        assert list(p["shift"]) == [0, 0, 0]
        assert p["distance"] == 2.0


def test_small_cell_large_cutoff():
    # neighbors.rs:505-527
    pairs = orc.neighbors(np.zeros((1, 3)), np.eye(3) * 0.5, 0.6)
    assert len(pairs) == 3
    assert sorted(tuple(p["shift"]) for p in pairs) == [(0, 0, 1), (0, 1, 0), (1, 0, 0)]


def test_non_cubic_cell():
    # neighbors.rs:529-561
    cell = np.array([[4.26, -2.45951215, 0.0], [2.13, 1.22975607, 0.0], [0.0, 0.0, 50.0]])
    pos = np.array([[1.42, 0, 0], [2.84, 0, 0], [3.55, -1.22975607, 0], [4.97, -1.22975607, 0]])
    pairs = orc.neighbors(pos, cell, 6.4)
    assert len(pairs) == 90
    found = {(int(p["first"]), int(p["second"]), tuple(int(x) for x in p["shift"])) for p in pairs}
    for missing in [(0, 3, (-2, 0, 0)), (0, 3, (-2, 1, 0)), (0, 3, (-2, 2, 0))]:
        assert missing in found


def brute_force(pos, cell, cutoff):
    """All (i <= j, shift) within the cutoff, same floating-point formula."""
    n = len(pos)
    periodic = np.any(cell != 0)
    out = []
    if periodic:
        inv = np.linalg.inv(cell)
        frac = pos @ inv
        faces = np.array([abs(np.dot(np.cross(cell[(k + 1) % 3], cell[(k + 2) % 3]) / np.linalg.norm(np.cross(cell[(k + 1) % 3], cell[(k + 2) % 3])), cell[k])) for k in range(3)])
        span = [int(math.ceil(cutoff / faces[k])) + 1 + int(np.ceil(np.max(np.abs(frac[:, k])))) for k in range(3)]
        shifts = list(itertools.product(*[range(-s, s + 1) for s in span]))
    else:
        shifts = [(0, 0, 0)]
    for i in range(n):
        for j in range(i, n):
            for s in shifts:
                if i == j:
                    if s == (0, 0, 0):
                        continue
                    t = s[0] + s[1] + s[2]
                    if t < 0 or (t == 0 and (s[2] < 0 or (s[2] == 0 and s[1] < 0))):
                        continue
                v = pos[j] - pos[i]
                sx = cell[0, 0] * s[0] + cell[1, 0] * s[1] + cell[2, 0] * s[2]
                sy = cell[0, 1] * s[0] + cell[1, 1] * s[1] + cell[2, 1] * s[2]
                sz = cell[0, 2] * s[0] + cell[1, 2] * s[1] + cell[2, 2] * s[2]
                v = np.array([v[0] + sx, v[1] + sy, v[2] + sz])
                d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
                if d2 < cutoff * cutoff:
                    out.append((i, j, s, v, math.sqrt(d2)))
    out.sort(key=lambda x: (x[0], x[1], x[2]))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_bit_exact(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 25))
    if seed % 3 == 0:
        cell = np.zeros((3, 3))
        pos = rng.uniform(0, 6, size=(n, 3))
    elif seed % 3 == 1:
        L = rng.uniform(3.0, 9.0)
        cell = np.eye(3) * L
        pos = rng.uniform(-2, L + 2, size=(n, 3))   # includes atoms outside the cell
    else:
        cell = np.array([[5.0, 0, 0], [1.3, 4.6, 0], [-0.8, 1.1, 5.2]]) * rng.uniform(0.6, 1.4)
        pos = rng.uniform(0, 5, size=(n, 3)) @ np.eye(3)
    cutoff = rng.uniform(2.0, 6.5)
    got = orc.neighbors(pos, cell, cutoff)
    exp = brute_force(pos, cell, cutoff)
    assert len(got) == len(exp)
    for p, (i, j, s, v, d) in zip(got, exp):
        assert (int(p["first"]), int(p["second"]), tuple(int(x) for x in p["shift"])) == (i, j, s)
        assert tuple(p["vector"]) == tuple(v)
        assert p["distance"] == d


# ---------------------------------------------------------------- NeighborList calculator
# Known answers of featomic/src/calculators/neighbor_list.rs:640-958, on the test
# systems of featomic/src/systems/test_utils.rs:29-43.

def water():
    return orc.System([[0.0, 0.0, 0.0], [0.0, 0.75545, -0.58895], [0.0, -0.75545, -0.58895]], [-42, 1, 1],
                      np.eye(3) * 3.0)


def ch():
    return orc.System([[0.0, 0.0, 0.0], [0.0, 1.2, 0.0]], [6, 1], np.eye(3) * 10.0)


def nl_block(tm, key):
    keys = [tuple(k) for k in np.asarray(tm["keys"]).tolist()]
    return tm["blocks"][keys.index(tuple(key))]


def test_neighbor_list_half_water():
    # neighbor_list.rs:640-692
    tm = orc.neighbor_list([water()], {"cutoff": 2.0, "full_neighbor_list": False, "self_pairs": False})
    assert np.asarray(tm["keys"]).tolist() == [[-42, 1], [1, 1]]
    oh = nl_block(tm, (-42, 1))
    assert oh["properties"].tolist() == [[1]]
    assert [c[0] for c in oh["components"]] == ["pair_xyz"]
    assert oh["samples"].tolist() == [[0, 0, 1, 0, 0, 0], [0, 0, 2, 0, 0, 0]]
    np.testing.assert_allclose(oh["values"][..., 0], [[0.0, 0.75545, -0.58895], [0.0, -0.75545, -0.58895]], rtol=1e-6)
    hh = nl_block(tm, (1, 1))
    assert hh["samples"].tolist() == [[0, 1, 2, 0, 0, 0], [0, 1, 2, 0, 1, 0]]
    np.testing.assert_allclose(hh["values"][..., 0], [[0.0, -1.5109, 0.0], [0.0, 1.4891, 0.0]], rtol=1e-6)


def test_neighbor_list_full_water():
    # neighbor_list.rs:694-770
    tm = orc.neighbor_list([water()], {"cutoff": 2.0, "full_neighbor_list": True, "self_pairs": False})
    assert np.asarray(tm["keys"]).tolist() == [[-42, 1], [1, -42], [1, 1]]
    oh = nl_block(tm, (-42, 1))
    assert oh["samples"].tolist() == [[0, 0, 1, 0, 0, 0], [0, 0, 2, 0, 0, 0]]
    np.testing.assert_allclose(oh["values"][..., 0], [[0.0, 0.75545, -0.58895], [0.0, -0.75545, -0.58895]], rtol=1e-6)
    ho = nl_block(tm, (1, -42))
    assert ho["samples"].tolist() == [[0, 1, 0, 0, 0, 0], [0, 2, 0, 0, 0, 0]]
    np.testing.assert_allclose(ho["values"][..., 0], [[0.0, -0.75545, 0.58895], [0.0, 0.75545, 0.58895]], rtol=1e-6)
    hh = nl_block(tm, (1, 1))
    assert hh["samples"].tolist() == [[0, 1, 2, 0, 0, 0], [0, 2, 1, 0, 0, 0], [0, 1, 2, 0, 1, 0], [0, 2, 1, 0, -1, 0]]
    np.testing.assert_allclose(hh["values"][..., 0],
                               [[0.0, -1.5109, 0.0], [0.0, 1.5109, 0.0], [0.0, 1.4891, 0.0], [0.0, -1.4891, 0.0]],
                               rtol=1e-6)


def test_neighbor_list_periodic_ch():
    # neighbor_list.rs:772-838: pairs of an atom with its own images
    half = orc.neighbor_list([ch()], {"cutoff": 12.0, "full_neighbor_list": False, "self_pairs": False})
    assert np.asarray(half["keys"]).tolist() == [[1, 1], [1, 6], [6, 6]]
    hh = nl_block(half, (1, 1))
    assert hh["samples"].tolist() == [[0, 1, 1, 0, 0, 1], [0, 1, 1, 0, 1, 0], [0, 1, 1, 1, 0, 0]]
    np.testing.assert_allclose(hh["values"][..., 0], [[0, 0, 10], [0, 10, 0], [10, 0, 0]], rtol=1e-6)
    full = orc.neighbor_list([ch()], {"cutoff": 12.0, "full_neighbor_list": True, "self_pairs": False})
    assert np.asarray(full["keys"]).tolist() == [[1, 1], [1, 6], [6, 1], [6, 6]]
    hh = nl_block(full, (1, 1))
    assert hh["samples"].tolist() == [[0, 1, 1, 0, 0, 1], [0, 1, 1, 0, 0, -1], [0, 1, 1, 0, 1, 0],
                                      [0, 1, 1, 0, -1, 0], [0, 1, 1, 1, 0, 0], [0, 1, 1, -1, 0, 0]]
    np.testing.assert_allclose(hh["values"][..., 0],
                               [[0, 0, 10], [0, 0, -10], [0, 10, 0], [0, -10, 0], [10, 0, 0], [-10, 0, 0]], rtol=1e-6)


def test_neighbor_list_self_pairs():
    # neighbor_list.rs:906-940
    tm = orc.neighbor_list([water()], {"cutoff": 2.0, "full_neighbor_list": True, "self_pairs": True})
    assert np.asarray(tm["keys"]).tolist() == [[-42, -42], [-42, 1], [1, -42], [1, 1]]
    assert tm["blocks"][3]["samples"].tolist() == [[0, 1, 2, 0, 0, 0], [0, 2, 1, 0, 0, 0], [0, 1, 2, 0, 1, 0],
                                                   [0, 2, 1, 0, -1, 0], [0, 1, 1, 0, 0, 0], [0, 2, 2, 0, 0, 0]]


def test_neighbor_list_empty():
    # neighbor_list.rs:942-958
    tm = orc.neighbor_list([water()], {"cutoff": 0.1, "full_neighbor_list": True, "self_pairs": False})
    assert np.asarray(tm["keys"]).shape == (0, 2)
    assert tm["blocks"] == []


@pytest.mark.parametrize("full", [False, True])
@pytest.mark.parametrize("system", ["water", "ch", "random"])
def test_neighbor_list_gradients_finite_differences(full, system):
    """Central differences of the pair vectors against the gradient rows (the
    reference's finite_differences_positions test, neighbor_list.rs:840-864, with a
    displacement of 1e-6); the CH system at cutoff 12 adds pairs of an atom with its
    own periodic images, whose derivative is zero."""
    if system == "water":
        sys_, cutoff = water(), 1.0
    elif system == "ch":
        sys_, cutoff = ch(), 12.0
    else:
        rng = np.random.default_rng(3)
        sys_ = orc.System(rng.uniform(0, 4.0, size=(7, 3)), rng.choice([1, 8], size=7), np.eye(3) * 4.0 + 0.3 * rng.normal(size=(3, 3)))
        cutoff = 3.1
    params = {"cutoff": cutoff, "full_neighbor_list": full, "self_pairs": True}
    ref = orc.neighbor_list([sys_], params, ["positions"])
    h = 1e-6
    for atom in range(len(sys_)):
        for x in range(3):
            moved = []
            for sign in (1, -1):
                pos = sys_.positions.copy()
                pos[atom, x] += sign * h
                moved.append(orc.neighbor_list([orc.System(pos, sys_.types, sys_.cell)], params))
            for b, blk in enumerate(ref["blocks"]):
                plus, minus = moved[0]["blocks"][b], moved[1]["blocks"][b]
                assert np.array_equal(plus["samples"], blk["samples"]) and np.array_equal(minus["samples"], blk["samples"])
                fd = (plus["values"] - minus["values"]) / (2 * h)
                analytic = np.zeros_like(blk["values"])
                g = blk["gradients"]["positions"]
                for r, (sample, _, a) in enumerate(g["samples"]):
                    if a == atom:
                        analytic[sample] += g["values"][r, x]
                np.testing.assert_allclose(analytic, fd, atol=1e-8)
