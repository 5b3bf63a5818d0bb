"""Regression test for the gradient-row kernels on centers without neighbours.

A center with no neighbour window still writes its own rows, and reads the LDS
table of present types to do so; a missing barrier once let such a center read a
stale table and index the self gradients out of bounds. These batches mix
isolated atoms (farther than the cutoff from every other atom) with dense
centers, and a system in which one of the batch's types is absent, and run the
power spectrum and the spherical expansion in FP64 and FP32 (plain and fast)
over every launch shape the autotuner can choose, against the FP64 oracle.
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import assert_block_approx, compare_tensormaps, oracle_to_tensormap, random_system

pytestmark = pytest.mark.gpu

HYPERS = {
    "cutoff": {"radius": 4.5, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 5, "radial": {"type": "Gto", "max_radial": 7}},
}
# FP32 floors of the element-wise bar (see test_gpu_workloads.py); the spherical
# expansion's rows are the FP32 spline derivatives themselves, summed over the
# neighbours for the center's own rows (3e-5 of the block's scale)
FLOOR = {False: 1e-5, True: 3e-5}
SE_FLOOR = 3e-5


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def systems():
    rng = np.random.default_rng(77)
    pos, types, _ = random_system(rng, 24, "cluster", 6.0, types=(1, 6, 8))
    # three isolated atoms, one of each type, more than 10 A from everything else
    isolated = np.array([[25.0, 25.0, 25.0], [25.0, -15.0, 3.0], [-14.0, 20.0, -12.0]])
    cluster = orc.System(np.concatenate([pos, isolated]), np.concatenate([types, [1, 6, 8]]).astype(np.int32),
                         np.zeros((3, 3)))
    # a single isolated atom alone in its system
    lone = orc.System(np.array([[0.0, 0.0, 0.0]]), np.array([8], dtype=np.int32), np.zeros((3, 3)))
    # a periodic cell without type 8
    pos, types, cell = random_system(rng, 30, "periodic", 7.0, types=(1, 6))
    periodic = orc.System(pos, types, cell)
    return [cluster, lone, periodic]


@pytest.fixture(scope="module")
def references():
    syst = systems()
    return (oracle_to_tensormap(orc.soap_power_spectrum(syst, HYPERS, ["positions"])),
            oracle_to_tensormap(orc.soap_spherical_expansion(syst, HYPERS, ["positions"])))


def check(tm, ref, precision, fast, floor):
    if precision == 64:
        compare_tensormaps(tm, ref, 1e-10, 1e-10, gradients=["positions"])
        return
    assert np.array_equal(tm.keys.values, ref.keys.values)
    for bg, be in zip(tm.blocks(), ref.blocks()):
        assert bg.samples == be.samples
        assert_block_approx(bg.values, be.values, 1e-4, floor, "values")
        gg, ge = bg.gradient("positions"), be.gradient("positions")
        assert gg.samples == ge.samples
        assert_block_approx(gg.values, ge.values, 1e-4, floor, "gradients")


@pytest.mark.parametrize("precision,fast", [(64, False), (32, False), (32, True)])
def test_isolated_centers_every_launch_shape(references, precision, fast):
    ps_ref, se_ref = references
    syst = [ra.System(s.types, s.positions, s.cell) for s in systems()]
    windows = list(range(13)) + [-2, -1]
    for kind, ref in (("power spectrum", ps_ref), ("spherical expansion", se_ref)):
        if kind == "spherical expansion" and fast:
            continue
        for w in windows:
            calc = (ra.SoapPowerSpectrum if kind == "power spectrum" else ra.SphericalExpansion)(**HYPERS)
            calc.set_precision(precision)
            calc.set_fast_form(fast)
            calc.tuning = {"expansion_entries": 0, "gradient_window": w}
            say(kind, "precision", precision, "fast", fast, "gradient window", w)
            tm = calc.compute(syst, gradients=["positions"])
            check(tm, ref, precision, fast, FLOOR[fast] if kind == "power spectrum" else SE_FLOOR)
