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
