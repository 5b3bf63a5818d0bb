"""The rank-2 gradient-row kernel (FP32 power spectrum, n = 8) on the cases its plan
and its loops branch on, against the FP64 oracle and against the tiled kernel:

* more than four present types (a second chunk of types, with two types in it);
* three present types (one lane group of four without a block);
* runs of several periodic images, and runs longer than a window (a cell of a few
  atoms, whose centers go to the tiled kernel);
* windows of eight single-image runs (the branch-free path) and partial windows;
* centers whose only neighbours are their own images.

Each launch shape that selects the rank-2 kernel (gradient windows 11 and 12, and
the built-in choice) is checked element-wise at 1e-4 with a floor of 1e-5 of the
block's largest value, and bit-for-bit labels.
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import assert_block_approx, oracle_to_tensormap, random_system

pytestmark = pytest.mark.gpu

HYPERS = {
    "cutoff": {"radius": 5.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 6, "radial": {"type": "Gto", "max_radial": 7}},
}
FLOOR = 1e-5
RANK2_WINDOWS = [11, 12, -1]


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def systems():
    rng = np.random.default_rng(2024)
    out = []
    # six types: two chunks of present types (4 + 2)
    pos, types, cell = random_system(rng, 40, "periodic", 8.0, types=(1, 6, 7, 8, 9, 16))
    out.append(orc.System(pos, types, cell))
    # three types in a cluster: lane groups without a block
    pos, types, cell = random_system(rng, 30, "cluster", 7.0, types=(1, 6, 8))
    out.append(orc.System(pos, types, cell))
    # a cell of three atoms: runs of many images, some longer than a window
    pos, types, cell = random_system(rng, 3, "periodic", 3.4, types=(1, 8))
    out.append(orc.System(pos, types, cell))
    # one atom in a small cell: its neighbours are its own images only
    out.append(orc.System(np.array([[0.3, 0.2, 0.1]]), np.array([6], dtype=np.int32), np.eye(3) * 3.0))
    # a dense triclinic cell of two types: full windows of single-image runs
    pos, types, cell = random_system(rng, 48, "triclinic", 8.0, types=(1, 8))
    out.append(orc.System(pos, types, cell))
    return out


@pytest.fixture(scope="module")
def reference():
    return oracle_to_tensormap(orc.soap_power_spectrum(systems(), HYPERS, ["positions"]))


def compute(window):
    syst = [ra.System(s.types, s.positions, s.cell) for s in systems()]
    calc = ra.SoapPowerSpectrum(**HYPERS)
    calc.set_precision(32)
    calc.tuning = {"expansion_entries": 0, "gradient_window": window}
    return calc.compute(syst, gradients=["positions"])


@pytest.mark.parametrize("window", RANK2_WINDOWS)
def test_rank2_against_oracle(reference, window):
    say("gradient window", window)
    tm = compute(window)
    assert np.array_equal(tm.keys.values, reference.keys.values)
    worst = 0.0
    for bg, be in zip(tm.blocks(), reference.blocks()):
        assert bg.samples == be.samples
        assert_block_approx(bg.values, be.values, 1e-4, FLOOR, "values")
        gg, ge = bg.gradient("positions"), be.gradient("positions")
        assert gg.samples == ge.samples
        worst = max(worst, assert_block_approx(gg.values, ge.values, 1e-4, FLOOR, "gradients"))
    say(f"  floor needed by the gradients: {worst:.2e}")


def test_rank2_against_tiled_kernel():
    """The rank-2 and tiled kernels on the same input agree to FP32 rounding."""
    rank2 = compute(11)
    tiled = compute(1)
    for b2, bt in zip(rank2.blocks(), tiled.blocks()):
        g2, gt = b2.gradient("positions").values, bt.gradient("positions").values
        assert g2.shape == gt.shape
        scale = max(np.abs(gt).max(), 1e-30)
        assert np.abs(g2 - gt).max() <= 2e-6 * scale
