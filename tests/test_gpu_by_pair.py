"""GPU tests of SphericalExpansionByPair (featomic/src/calculators/soap/spherical_expansion_pair.rs):
one sample per directed pair, written by by_pair_kernel.

* against the FP64 oracle (which the reference's own sums-to-expansion and
  finite-difference checks pin): keys, samples, gradient samples identical; values,
  position, strain and cell gradients within |a - b| <= 1e-10 (|b| + max|b|) in FP64,
  and block-relative 1e-4 in FP32; on the reference's water + methane systems with
  its test parameters (spherical_expansion_pair.rs:861-887) and on a random batch
  (periodic, triclinic, cluster; exact radial integral);
* the reference's sums_to_spherical_expansion test (spherical_expansion_pair.rs:
  1000-1040) against this library's own spherical expansion;
* use_native_system = False (the caller's pairs) gives the same TensorMap.
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import block_relative_error, compare_tensormaps, oracle_to_tensormap, random_system

pytestmark = pytest.mark.gpu

PARAMS = {
    "cutoff": {"radius": 7.3, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3,
                "scaling": {"type": "Willatt2018", "scale": 1.5, "rate": 0.8, "exponent": 2.0}},
    "basis": {"type": "TensorProduct", "max_angular": 2, "radial": {"type": "Gto", "max_radial": 5},
              "spline_accuracy": 1e-8},
}
RANDOM = {
    "cutoff": {"radius": 4.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5},
              "spline_accuracy": None},
}
GRADS = ["positions", "strain", "cell"]


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def reference_systems():
    water = (np.array([[0.0, 0.0, 0.0], [0.0, 0.75545, -0.58895], [0.0, -0.75545, -0.58895]]),
             np.array([-42, 1, 1], dtype=np.int32), np.eye(3) * 3.0)
    methane = (np.array([[5.0, 5.0, 5.0], [5.5288, 5.1610, 5.9359], [5.2051, 5.8240, 4.3214],
                         [5.3345, 4.0686, 4.5504], [3.9315, 4.9463, 5.1921]]),
               np.array([6, 1, 1, 1, 1], dtype=np.int32), np.eye(3) * 5.0)
    return [water, methane]


def random_batch(seed):
    rng = np.random.default_rng(seed)
    return [random_system(rng, 14, "periodic", 6.5), random_system(rng, 11, "triclinic", 6.0, types=(1, 8)),
            random_system(rng, 12, "cluster", 8.0)]


def oracle(systems, hypers, gradients):
    return oracle_to_tensormap(orc.soap_spherical_expansion_by_pair([orc.System(*s) for s in systems], hypers,
                                                                    gradients))


def compute(systems, hypers, precision, gradients, native=True):
    calc = ra.SphericalExpansionByPair(**hypers)
    calc.set_precision(precision)
    if native:
        ra_sys = [ra.System(t, p, c) for p, t, c in systems]
    else:
        rc = hypers["cutoff"]["radius"]
        ra_sys = [ra.System(t, p, c, pairs=ra.neighbors(p, c, rc)) for p, t, c in systems]
    return calc.compute(ra_sys, gradients=gradients, use_native_system=native)


@pytest.mark.parametrize("case", ["reference", "random"])
def test_by_pair_fp64_vs_oracle(case):
    say("by pair FP64 vs oracle:", case)
    systems, hypers = (reference_systems(), PARAMS) if case == "reference" else (random_batch(21), RANDOM)
    got = compute(systems, hypers, 64, GRADS)
    ref = oracle(systems, hypers, GRADS)
    worst = compare_tensormaps(got, ref, 1e-10, 1e-10, gradients=GRADS)
    say(f"  worst ratio to the bound {worst:.2e}")


@pytest.mark.parametrize("case", ["reference", "random"])
def test_by_pair_fp32_vs_oracle(case):
    say("by pair FP32 vs oracle:", case)
    systems, hypers = (reference_systems(), PARAMS) if case == "reference" else (random_batch(22), RANDOM)
    got = compute(systems, hypers, 32, GRADS)
    ref = oracle(systems, hypers, GRADS)
    compare_tensormaps(got, ref, 1.0, 1.0, gradients=GRADS)  # labels
    for param in [None] + GRADS:
        err = block_relative_error(got, ref, param)
        say(f"  {param or 'values'}: block-relative error {err:.2e}")
        assert err <= 1e-4


def test_pairs_sum_to_spherical_expansion():
    say("by pair sums to the GPU spherical expansion")
    systems = reference_systems() + random_batch(23)[:1]
    by_pair = compute(systems, PARAMS, 64, [])
    calc = ra.SphericalExpansion(**PARAMS)
    calc.set_precision(64)
    se = calc.compute([ra.System(t, p, c) for p, t, c in systems])
    se_keys = [tuple(k) for k in np.asarray(se.keys.values).tolist()]
    pair_keys = [tuple(k) for k in np.asarray(by_pair.keys.values).tolist()]
    assert sorted(se_keys) == sorted(pair_keys)
    # on the scale of the whole expansion: blocks that vanish by symmetry (l = 1 of an
    # atom's own images) hold rounding residues of ~1e-18 in the expansion and exact
    # zeros in the pair sums
    scale = max(float(np.abs(np.asarray(b.values)).max()) for b in se.blocks() if np.asarray(b.values).size)
    for key, blk in zip(se_keys, se.blocks()):
        pb = by_pair.block(pair_keys.index(key))
        index = {(int(s), int(a)): k for k, (s, a) in enumerate(np.asarray(blk.samples.values))}
        summed = np.zeros_like(np.asarray(blk.values))
        for k, smp in enumerate(np.asarray(pb.samples.values)):
            summed[index[(int(smp[0]), int(smp[1]))]] += np.asarray(pb.values)[k]
        ref = np.asarray(blk.values)
        assert np.abs(summed - ref).max() <= 1e-12 * scale, key


def test_by_pair_user_pairs():
    say("by pair with the caller's pairs")
    systems = random_batch(24)
    a = compute(systems, RANDOM, 64, ["positions"], native=False)
    b = compute(systems, RANDOM, 64, ["positions"], native=True)
    compare_tensormaps(a, b, 1e-14, 1e-14, gradients=["positions"])


def test_by_pair_isolated_atoms_and_empty_system():
    """Atoms without neighbours: the l = 0 blocks hold only their self pairs and the
    l > 0 blocks of those types are empty; a system without atoms adds nothing."""
    say("by pair: isolated atoms, empty system")
    isolated = (np.array([[0.0, 0.0, 0.0], [20.0, 0.0, 0.0]]), np.array([1, 8], dtype=np.int32), np.zeros((3, 3)))
    empty = (np.zeros((0, 3)), np.zeros(0, dtype=np.int32), np.zeros((3, 3)))
    systems = [isolated, empty] + random_batch(25)[:1]
    got = compute(systems, RANDOM, 64, GRADS)
    ref = oracle(systems, RANDOM, GRADS)
    compare_tensormaps(got, ref, 1e-10, 1e-10, gradients=GRADS)
    only = compute([isolated], RANDOM, 64, ["positions"])
    assert [tuple(k) for k in np.asarray(only.keys.values).tolist()] == [(l, 1, t, t) for t in (1, 8) for l in range(5)]
    for k, blk in zip(np.asarray(only.keys.values), only.blocks()):
        assert len(blk.samples) == (1 if k[0] == 0 else 0)
        assert len(blk.gradient("positions").samples) == 0
