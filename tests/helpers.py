"""Shared test helpers: fixture loading, oracle dict -> TensorMap, comparisons."""

import gzip
import os

import numpy as np

from rascaline_amd.tensormap import Labels, TensorBlock, TensorMap

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return os.path.join(GOLDEN, name)


def load_npy_gz(name):
    with gzip.open(golden(name)) as fd:
        return np.load(fd)


def _labels(names, values):
    return Labels(names, values)


def _block_from_dict(d):
    comps = [Labels([name], values) for (name, values) in d["components"]]
    blk = TensorBlock(d["values"], _labels(d["samples_names"], d["samples"]), comps,
                      _labels(d["properties_names"], d["properties"]))
    for param, g in d["gradients"].items():
        blk.add_gradient(param, _block_from_dict(g))
    return blk


def oracle_to_tensormap(d):
    return TensorMap(Labels(d["keys_names"], d["keys"]), [_block_from_dict(b) for b in d["blocks"]])


def flatten_se(tm):
    """Flattening used by featomic/tests/soap-spherical-expansion.rs:17-24."""
    tm = tm.keys_to_samples("center_type")
    tm = tm.keys_to_properties("neighbor_type")
    tm = tm.components_to_properties("o3_mu")
    tm = tm.keys_to_properties("o3_lambda")
    return tm


def flatten_ps(tm):
    """Flattening used by featomic/tests/soap-power-spectrum.rs:17-21."""
    tm = tm.keys_to_samples("center_type")
    tm = tm.keys_to_properties(["neighbor_1_type", "neighbor_2_type"])
    return tm


def sum_gradients(n_atoms, gradient):
    """featomic/tests/soap-spherical-expansion.rs:102-114."""
    out = np.zeros((n_atoms, 3, gradient.values.shape[-1]))
    atoms = gradient.samples.column("atom")
    np.add.at(out, atoms, gradient.values.reshape(len(atoms), 3, -1))
    return out


def assert_relative_eq(actual, expected, max_relative, epsilon=0.0):
    """approx::assert_relative_eq semantics: |a-b| <= eps or |a-b| <= rel * max(|a|, |b|)."""
    actual = np.asarray(actual)
    expected = np.asarray(expected)
    assert actual.shape == expected.shape, (actual.shape, expected.shape)
    diff = np.abs(actual - expected)
    bound = np.maximum(epsilon, max_relative * np.maximum(np.abs(actual), np.abs(expected)))
    bad = diff > bound
    if np.any(bad):
        idx = np.argwhere(bad)[:5]
        details = [(tuple(i), actual[tuple(i)], expected[tuple(i)]) for i in idx]
        raise AssertionError(f"{int(bad.sum())} / {bad.size} elements differ, worst rel "
                             f"{float(np.max(diff / np.maximum(np.abs(expected), 1e-300)))}: {details}")


def max_rel_err(actual, expected, floor=0.0):
    actual = np.asarray(actual)
    expected = np.asarray(expected)
    denom = np.maximum(np.maximum(np.abs(actual), np.abs(expected)), floor)
    denom = np.where(denom == 0, 1.0, denom)
    return float(np.max(np.abs(actual - expected) / denom)) if actual.size else 0.0


def random_system(rng, n_atoms, kind="periodic", L=8.0, types=(1, 6, 8), min_dist=0.8):
    """Random atoms with a minimum-image rejection distance.

    kind: "periodic" (cubic L), "triclinic", "cluster" (no cell), "outside"
    (cubic, atoms placed outside the unit cell too)."""
    if kind == "triclinic":
        cell = np.array([[L, 0, 0], [0.3 * L, 0.9 * L, 0], [-0.2 * L, 0.25 * L, 0.95 * L]])
    elif kind == "cluster":
        cell = np.zeros((3, 3))
    else:
        cell = np.eye(3) * L
    pos = []
    tries = 0
    while len(pos) < n_atoms and tries < 100000:
        tries += 1
        if kind == "cluster":
            p = rng.uniform(0, L, 3)
        else:
            f = rng.uniform(0, 1, 3)
            p = f @ cell
            if kind == "outside":
                p = p + rng.integers(-2, 3, 3) * L
        ok = True
        for q in pos:
            d = p - q
            if kind != "cluster":
                frac = np.linalg.solve(cell.T, d)
                frac -= np.round(frac)
                d = frac @ cell
            if np.dot(d, d) < min_dist ** 2:
                ok = False
                break
        if ok:
            pos.append(p)
    pos = np.array(pos)
    t = rng.choice(np.array(types), size=len(pos))
    return pos, t.astype(np.int32), cell


def compare_tensormaps(actual, expected, rtol, atol_scale, gradients=(), check_labels=True):
    """Block-by-block comparison: identical keys and labels, and values within
    |a - b| <= rtol * |b| + atol_scale * S, where S is the largest |b| of the same
    array kind (values, or one gradient parameter) over the whole TensorMap, so that
    elements that vanish by symmetry or cancellation are compared on the scale of
    the descriptor rather than on their own."""
    assert actual.keys.names == expected.keys.names
    assert np.array_equal(actual.keys.values, expected.keys.values), (actual.keys.values, expected.keys.values)

    def scale_of(get):
        m = 0.0
        for blk in expected.blocks():
            v = np.asarray(get(blk))
            if v.size:
                m = max(m, float(np.max(np.abs(v))))
        return m

    value_scale = scale_of(lambda b: b.values)
    grad_scale = {p: scale_of(lambda b, p=p: b.gradient(p).values) for p in gradients}
    worst = 0.0
    for ba, be in zip(actual.blocks(), expected.blocks()):
        if check_labels:
            assert ba.samples == be.samples, "samples differ"
            assert ba.properties == be.properties, "properties differ"
            assert len(ba.components) == len(be.components)
            for ca, ce in zip(ba.components, be.components):
                assert ca == ce
        worst = max(worst, _close(ba.values, be.values, rtol, atol_scale * value_scale))
        for param in gradients:
            ga, ge = ba.gradient(param), be.gradient(param)
            if check_labels:
                assert ga.samples == ge.samples, f"{param} gradient samples differ"
                for ca, ce in zip(ga.components, ge.components):
                    assert ca == ce
            worst = max(worst, _close(ga.values, ge.values, rtol, atol_scale * grad_scale[param]))
    return worst


def _close(a, b, rtol, atol):
    """Returns the largest |a - b| / (|b| + atol / rtol) ratio after checking the bound."""
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if b.size == 0:
        return 0.0
    diff = np.abs(a - b)
    bound = rtol * np.abs(b) + atol
    bad = diff > bound
    if np.any(bad):
        idx = np.argwhere(bad)[:5]
        raise AssertionError(f"{int(bad.sum())}/{bad.size} elements differ (atol {atol:.3e}): " +
                             ", ".join(f"{tuple(i)}: {a[tuple(i)]!r} vs {b[tuple(i)]!r}" for i in idx))
    return float(np.max(diff / np.maximum(np.abs(b) + atol / max(rtol, 1e-300), 1e-300)))


def block_relative_error(actual, expected, param=None):
    """Largest, over blocks, of max|a - b| / max|b| within the block (values, or one
    gradient parameter). Element-wise relative errors are not a usable bar for FP32:
    elements that cancel to ~1e-6 of their block's scale carry relative errors of
    order 1e-2 in plain FP32 arithmetic already."""
    get = (lambda b: b.values) if param is None else (lambda b: b.gradient(param).values)
    worst = 0.0
    for ba, be in zip(actual.blocks(), expected.blocks()):
        a, b = np.asarray(get(ba)), np.asarray(get(be))
        if not b.size:
            continue
        scale = float(np.abs(b).max())
        if scale > 0:
            worst = max(worst, float(np.abs(a - b).max()) / scale)
    return worst


def approx_floor_needed(actual, expected, rel=1e-4, scale=None):
    """Smallest eps_frac for which every element of a block passes the reference tests'
    approx semantics, |a - b| <= eps or |a - b| <= rel * max(|a|, |b|), with the absolute
    floor eps = eps_frac * max|expected| of the block (or eps_frac * scale when the arrays
    are a part of a block whose largest element is `scale`)."""
    a = np.asarray(actual, dtype=np.float64)
    b = np.asarray(expected, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    if b.size == 0:
        return 0.0
    if scale is None:
        scale = float(np.abs(b).max())
    diff = np.abs(a - b)
    fail = diff > rel * np.maximum(np.abs(a), np.abs(b))
    if not np.any(fail):
        return 0.0
    if scale == 0.0:
        return float("inf")
    return float(diff[fail].max() / scale)


def assert_block_approx(actual, expected, rel, eps_frac, what="", scale=None):
    """Element-wise bar: |a - b| <= max(eps_frac * max|expected|, rel * max(|a|, |b|)) over one
    block's array; returns the floor the block needed (see approx_floor_needed)."""
    needed = approx_floor_needed(actual, expected, rel, scale)
    assert needed <= eps_frac, f"{what}: needs eps {needed:.2e} x max|ref| (bar {eps_frac:.0e}) at rel {rel:.0e}"
    return needed
