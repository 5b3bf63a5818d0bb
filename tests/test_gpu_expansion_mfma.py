"""The matrix-core expansion kernel (FP32, n = 8, l_max < 9) on the cases its chunks,
type groups and block layout branch on, against the FP64 oracle and against the
vector-ALU expansion kernel:

* more than four types (two passes over the entries, one per group of four types);
* three types (a group with an empty type);
* a cell of a few atoms, whose centers see their own periodic images (values only,
  no self gradient), and one atom whose only neighbours are its own images;
* an isolated atom of a cluster (no entries: the self contribution alone);
* l_max 0, 1, 2, 6, 7 and 8 (one to four halves of 16 blocks per instruction);
* chunks of 8 and 16 entries, with centers of more entries than a chunk.

The spherical expansion's values and position gradients are checked element-wise at
1e-4 with a floor of 1e-5 of the block's largest value; the power spectrum built on
them likewise. The fast form's values (bfloat16 hi + lo products in the same kernel) are
checked against the oracle at 1e-4 of each block's largest value.
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import assert_block_approx, oracle_to_tensormap, random_system

pytestmark = pytest.mark.gpu

FLOOR = 1e-5


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def hypers(lmax):
    return {
        "cutoff": {"radius": 5.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
        "density": {"type": "Gaussian", "width": 0.3},
        "basis": {"type": "TensorProduct", "max_angular": lmax, "radial": {"type": "Gto", "max_radial": 7}},
    }


def systems():
    rng = np.random.default_rng(77)
    out = []
    # six types: two groups of four types
    pos, types, cell = random_system(rng, 40, "periodic", 8.0, types=(1, 6, 7, 8, 9, 16))
    out.append(orc.System(pos, types, cell))
    # three types in a cluster, with one atom far from the others
    pos, types, cell = random_system(rng, 24, "cluster", 7.0, types=(1, 6, 8))
    pos = np.vstack([pos, [[40.0, 40.0, 40.0]]])
    types = np.concatenate([types, np.array([6], dtype=types.dtype)])
    out.append(orc.System(pos, types, cell))
    # a cell of three atoms: each center sees its own images
    pos, types, cell = random_system(rng, 3, "periodic", 3.4, types=(1, 8))
    out.append(orc.System(pos, types, cell))
    # one atom in a small cell: its neighbours are its own images only
    out.append(orc.System(np.array([[0.3, 0.2, 0.1]]), np.array([6], dtype=np.int32), np.eye(3) * 3.0))
    return out


def compute(calc_class, lmax, entries):
    syst = [ra.System(s.types, s.positions, s.cell) for s in systems()]
    calc = calc_class(**hypers(lmax))
    calc.set_precision(32)
    calc.tuning = {"expansion_entries": entries, "gradient_window": -1}
    return calc.compute(syst, gradients=["positions"])


def check_against(tm, reference):
    assert np.array_equal(tm.keys.values, reference.keys.values)
    worst = 0.0
    for bg, be in zip(tm.blocks(), reference.blocks()):
        assert bg.samples == be.samples
        assert_block_approx(bg.values, be.values, 1e-4, FLOOR, "values")
        gg, ge = bg.gradient("positions"), be.gradient("positions")
        assert gg.samples == ge.samples
        worst = max(worst, assert_block_approx(gg.values, ge.values, 1e-4, FLOOR, "gradients"))
    return worst


@pytest.mark.parametrize("lmax", [0, 1, 2, 6, 7, 8])
@pytest.mark.parametrize("entries", [-8, -16])
def test_expansion_mfma_against_oracle(lmax, entries):
    say("spherical expansion, l_max", lmax, "entries", entries)
    reference = oracle_to_tensormap(orc.soap_spherical_expansion(systems(), hypers(lmax), ["positions"]))
    worst = check_against(compute(ra.SphericalExpansion, lmax, entries), reference)
    say(f"  floor needed by the gradients: {worst:.2e}")


@pytest.mark.parametrize("entries", [-8, -16])
def test_power_spectrum_on_mfma_expansion(entries):
    say("power spectrum, l_max 6, entries", entries)
    reference = oracle_to_tensormap(orc.soap_power_spectrum(systems(), hypers(6), ["positions"]))
    check_against(compute(ra.SoapPowerSpectrum, 6, entries), reference)


def test_expansion_mfma_against_vector_alu_kernel():
    """Both expansion kernels on the same input agree to FP32 rounding."""
    mfma = compute(ra.SphericalExpansion, 6, -8)
    valu = compute(ra.SphericalExpansion, 6, 8)
    for bm, bv in zip(mfma.blocks(), valu.blocks()):
        scale = max(np.abs(bv.values).max(), 1e-30)
        assert np.abs(bm.values - bv.values).max() <= 2e-6 * scale
        gm, gv = bm.gradient("positions").values, bv.gradient("positions").values
        assert gm.shape == gv.shape
        scale = max(np.abs(gv).max(), 1e-30)
        assert np.abs(gm - gv).max() <= 2e-6 * scale


@pytest.mark.parametrize("lmax", [6, 8])
def test_power_spectrum_values_bf16_split_in_expansion_kernel(lmax):
    """The fast form's values (products over m in bfloat16 hi + lo, formed in the expansion
    kernel; at l = 8 the 17 values of m take two products of 16) stay within 1e-4 of the FP64
    oracle per block, and agree with the FP32 products."""
    say("power spectrum values, bfloat16 hi + lo in the expansion kernel, l_max", lmax)
    reference = oracle_to_tensormap(orc.soap_power_spectrum(systems(), hypers(lmax)))
    syst = [ra.System(s.types, s.positions, s.cell) for s in systems()]
    out = {}
    for fast in (False, True):
        calc = ra.SoapPowerSpectrum(**hypers(lmax))
        calc.set_precision(32)
        calc.set_fast_form(fast)
        calc.tuning = {"expansion_entries": -8, "gradient_window": -1}
        out[fast] = calc.compute(syst)
    worst = 0.0
    for bf, bp, be in zip(out[True].blocks(), out[False].blocks(), reference.blocks()):
        scale = max(np.abs(be.values).max(), 1e-30)
        err = np.abs(bf.values - be.values).max() / scale
        worst = max(worst, err)
        assert err <= 1e-4
        assert np.abs(bf.values - bp.values).max() <= 1e-4 * scale
    say(f"  block-relative error of the split products: {worst:.2e}")
