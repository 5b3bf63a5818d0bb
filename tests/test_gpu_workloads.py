"""GPU parity on the benchmark workloads themselves, built by the seeded generators
of rascaline_amd.datasets (the same inputs bench.py times):

* 32 periodic cubic cells of 80-120 atoms, three types (the FP64 CPU-runnable
  workload): every cell's power spectrum against the FP64 oracle at 1e-10, the
  position gradients of six cells, and FP32 plain and fast forms element-wise;
* 4096 uneven structures (the headline workload): the first 64 (clusters, cubic
  and triclinic cells, 8-256 atoms) through the chunked two-stream device path in
  FP32 plain and fast against the FP64 oracle element-wise; then all 4096 in the
  benchmark's chunks of 56 000 atoms, checked by translation invariance of every
  gradient row set and fast against plain element by element;
* two LODE cells of 500-1500 atoms at the LODE workload's hyper-parameters, FP32
  with FP64 phases: every value, and the gradient rows of four centers per cell
  against the oracle, every center by translation invariance.

The element-wise bar is the reference tests' approx semantics with the metric's
relative tolerance: |a - b| <= max(eps, 1e-4 max(|a|, |b|)), with the absolute
floor eps = FLOOR x max|ref| of the block (values and gradients separately). The
floors are what FP32 arithmetic itself gives: the plain form's coefficients, their
derivatives and the contraction each round to FP32 and the gradient rows cancel
terms of the block's scale (block-relative error up to ~7e-6, floor 1e-5); the fast
form's bfloat16 hi + lo products drop lo*lo and the split residual, ~2^-16 of each
product on terms of the block's scale (floor 3e-5).
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from helpers import assert_block_approx, compare_tensormaps, oracle_to_tensormap
from rascaline_amd.datasets import batched_uneven, lode_cells, periodic_cubic_cells
from rascaline_amd.device import DeviceSystems, compute_device

pytestmark = pytest.mark.gpu

REL = 1e-4
FLOOR = {"plain": 1e-5, "fast": 3e-5}

SOAP = {
    "cutoff": {"radius": 5.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 6, "radial": {"type": "Gto", "max_radial": 7}},
}
LODE = {
    "density": {"type": "SmearedPowerLaw", "smearing": 1.0, "exponent": 1},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5, "radius": 5.0}},
}


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def ra_systems(raw):
    return [ra.System(t, p, c) for (p, t, c) in raw]


def orc_systems(raw):
    return [orc.System(p, t, c) for (p, t, c) in raw]


def power_spectrum(precision, fast=False):
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(precision)
    calc.set_fast_form(fast)
    return calc


# ---------------------------------------------------------------------------- 32 cubic cells, FP64


@pytest.fixture(scope="module")
def cubic_cells():
    raw = periodic_cubic_cells(seed=0)
    assert len(raw) == 32
    return raw


def test_cubic_cells_values_fp64(cubic_cells):
    say("32 cubic cells: power spectrum values, FP64")
    got = power_spectrum(64).compute(ra_systems(cubic_cells))
    ref = oracle_to_tensormap(orc.soap_power_spectrum(orc_systems(cubic_cells), SOAP))
    worst = compare_tensormaps(got, ref, 1e-10, 1e-10)
    say(f"  worst {worst:.1e}")


def test_cubic_cells_gradients_fp64(cubic_cells):
    say("cubic cells 0-5: power spectrum + position gradients, FP64")
    raw = cubic_cells[:6]
    got = power_spectrum(64).compute(ra_systems(raw), gradients=["positions"])
    ref = oracle_to_tensormap(orc.soap_power_spectrum(orc_systems(raw), SOAP, ["positions"]))
    worst = compare_tensormaps(got, ref, 1e-10, 1e-10, gradients=["positions"])
    say(f"  worst {worst:.1e}")


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_cubic_cells_fp32_elementwise(cubic_cells, form):
    say("cubic cells 0-3: FP32", form)
    raw = cubic_cells[:4]
    got = power_spectrum(32, form == "fast").compute(ra_systems(raw), gradients=["positions"])
    ref = oracle_to_tensormap(orc.soap_power_spectrum(orc_systems(raw), SOAP, ["positions"]))
    needed = [0.0, 0.0]
    for key, bg, be in zip(ref.keys.values, got.blocks(), ref.blocks()):
        assert bg.samples == be.samples
        needed[0] = max(needed[0], assert_block_approx(bg.values, be.values, REL, FLOOR[form], f"values {key}"))
        gg, ge = bg.gradient("positions"), be.gradient("positions")
        assert gg.samples == ge.samples
        needed[1] = max(needed[1], assert_block_approx(gg.values, ge.values, REL, FLOOR[form], f"gradients {key}"))
    say(f"  floors needed: values {needed[0]:.1e}, gradients {needed[1]:.1e}")


# ---------------------------------------------------------------------------- 4096 uneven structures


def device_blocks_vs_oracle(dev, ref, form):
    """Keys, labels, values and gradient rows of a device result against the oracle's dict."""
    import torch

    torch.cuda.synchronize()
    assert np.array_equal(dev.keys, ref["keys"]), (dev.keys, ref["keys"])
    needed = [0.0, 0.0]
    for b, rb in enumerate(ref["blocks"]):
        assert np.array_equal(dev.sample_labels(b).cpu().numpy(), rb["samples"])
        rg = rb["gradients"]["positions"]
        assert np.array_equal(dev.gradient_labels(b).cpu().numpy(), rg["samples"])
        v = dev.values(b).double().cpu().numpy()
        needed[0] = max(needed[0], assert_block_approx(v, rb["values"].reshape(v.shape), REL, FLOOR[form],
                                                       f"values {rb['samples'].shape} block {b}"))
        g = dev.gradients(b).double().cpu().numpy()
        needed[1] = max(needed[1], assert_block_approx(g, rg["values"].reshape(g.shape), REL, FLOOR[form],
                                                       f"gradients block {b}"))
    return needed


@pytest.fixture(scope="module")
def uneven_slice():
    raw = batched_uneven(seed=0, n_structures=64)
    chunks = [raw[:32], raw[32:]]
    refs = [orc.soap_power_spectrum(orc_systems(c), SOAP, ["positions"]) for c in chunks]
    return chunks, refs


@pytest.mark.parametrize("form", ["plain", "fast"])
def test_uneven_structures_slice_device_fp32(uneven_slice, form):
    """The first 64 structures of the headline workload in two chunks on two calculators
    with their own streams (the benchmark's pipelining), FP32, against the FP64 oracle."""
    import torch

    chunks, refs = uneven_slice
    say("first 64 of the 4096 uneven structures, device path, FP32", form,
        [max(len(t) for _, t, _ in c) for c in chunks])
    kinds = {("cluster" if not np.any(c) else "periodic") for c_ in chunks for (_, _, c) in c_}
    assert kinds == {"cluster", "periodic"}
    calcs = [power_spectrum(32, form == "fast") for _ in chunks]
    streams = [torch.cuda.Stream() for _ in chunks]
    torch.cuda.synchronize()
    devs = [compute_device(calc, DeviceSystems(c), ["positions"], stream=s) for calc, c, s in zip(calcs, chunks, streams)]
    torch.cuda.synchronize()
    for dev, ref in zip(devs, refs):
        needed = device_blocks_vs_oracle(dev, ref, form)
        say(f"  floors needed: values {needed[0]:.1e}, gradients {needed[1]:.1e}")


def make_chunks(raw, max_atoms):
    chunks, current, atoms = [], [], 0
    for s in raw:
        if current and atoms + len(s[1]) > max_atoms:
            chunks.append(current)
            current, atoms = [], 0
        current.append(s)
        atoms += len(s[1])
    if current:
        chunks.append(current)
    return chunks


def translation_residual(dev, b):
    """max over the block's samples of |sum over the sample's rows of dX/dr| / max|dX/dr|:
    the sum over all atoms of a descriptor's position gradients vanishes."""
    import torch

    g = dev.gradients(b)
    if g.shape[0] == 0:
        return 0.0
    rows = dev.gradient_labels(b)[:, 0].long()
    resid = torch.zeros((dev.n_samples[b],) + tuple(g.shape[1:]), dtype=torch.float64, device=g.device)
    step = 1 << 17
    for r0 in range(0, g.shape[0], step):
        resid.index_add_(0, rows[r0:r0 + step], g[r0:r0 + step].double())
    return float(resid.abs().max() / g.abs().max())


def fast_vs_plain_floor(dp, df, b):
    """Smallest floor (x max|plain|) for which |fast - plain| <= max(floor, 1e-4 max(|fast|, |plain|))
    over the block's gradient rows, computed on the device in row slices."""
    import torch

    gp, gf = dp.gradients(b), df.gradients(b)
    if gp.shape[0] == 0:
        return 0.0
    scale = float(gp.abs().max())
    worst = 0.0
    step = 1 << 18
    for r0 in range(0, gp.shape[0], step):
        p, f = gp[r0:r0 + step], gf[r0:r0 + step]
        diff = (f - p).abs()
        over = diff - REL * torch.maximum(f.abs(), p.abs())
        worst = max(worst, float(torch.where(over > 0, diff, torch.zeros_like(diff)).max()))
    return worst / scale if scale > 0 else 0.0


def test_uneven_structures_full_size():
    """All 4096 structures in the benchmark's chunks: every block's rows sum to zero per
    sample (translation invariance) in plain and fast FP32, and the fast form matches the
    plain one element-wise on every gradient row."""
    import torch

    raw = batched_uneven(seed=0)
    chunks = make_chunks(raw, 56000)
    say("4096 uneven structures, full size:", len(chunks), "chunks,", sum(len(t) for _, t, _ in raw), "atoms")
    plain, fast = power_spectrum(32), power_spectrum(32, True)
    worst = {"plain": 0.0, "fast": 0.0, "fast_vs_plain": 0.0}
    rows = 0
    for k, c in enumerate(chunks):
        ds = DeviceSystems(c)
        dp = compute_device(plain, ds, ["positions"])
        df = compute_device(fast, ds, ["positions"])
        torch.cuda.synchronize()
        assert np.array_equal(dp.keys, df.keys)
        for b in range(dp.n_blocks):
            assert dp.n_gradient_rows[b] == df.n_gradient_rows[b]
            rows += dp.n_gradient_rows[b]
            worst["plain"] = max(worst["plain"], translation_residual(dp, b))
            worst["fast"] = max(worst["fast"], translation_residual(df, b))
            worst["fast_vs_plain"] = max(worst["fast_vs_plain"], fast_vs_plain_floor(dp, df, b))
        say(f"  chunk {k}: {dp.total_rows} rows; worst so far {worst}")
    say(f"  {rows} gradient rows checked")
    assert rows > 70_000_000
    # translation residual of FP32 sums over up to a few hundred rows of the block's scale
    assert worst["plain"] <= 1e-5
    assert worst["fast"] <= 1e-4
    assert worst["fast_vs_plain"] <= FLOOR["fast"]


# ---------------------------------------------------------------------------- LODE cells

# LODE sums thousands of k-vector terms per element: FP32 floor 1e-5 of the block's scale
LODE_FLOOR = 1e-5


def lode_row_layout(types, ct, nt):
    """Samples (atoms of type ct) of block (ct, nt) of one system and the first gradient row
    of each: a sample's rows are the atoms j of type nt or j = i, in atom order
    (labels/samples/long_range.rs:55-77)."""
    samples = np.nonzero(types == ct)[0]
    per = np.count_nonzero(types == nt) + (types[samples] != nt).astype(np.int64)
    return samples, per, np.concatenate([[0], np.cumsum(per)])


@pytest.mark.parametrize("cell", [0, 1])
def test_lode_cells_fp32(cell):
    import torch

    raw = lode_cells(seed=0, n_structures=2)[cell]
    pos, types, box = raw
    n = len(types)
    say("LODE cell", cell, "of the 256-cell workload:", n, "atoms, FP32 with FP64 phases")
    calc = ra.LodeSphericalExpansion(**LODE)
    calc.set_precision(32)
    dev = compute_device(calc, DeviceSystems([raw]), ["positions"])
    torch.cuda.synchronize()
    species = np.unique(types)
    centers = [0, 1, n // 2, n - 1]
    values, grads, self_grads = orc.lode_dense(orc.System(pos, types, box), LODE, species, centers)
    N = LODE["basis"]["radial"]["max_radial"] + 1
    needed = [0.0, 0.0]
    worst_resid = 0.0
    for b, (l, _, ct, nt) in enumerate(dev.keys):
        ti = int(np.searchsorted(species, nt))
        sl = slice(l * l, (l + 1) * (l + 1))
        samples, per, offsets = lode_row_layout(types, ct, nt)
        v = dev.values(b).double().cpu().numpy().reshape(len(samples), 2 * l + 1, N)
        needed[0] = max(needed[0], assert_block_approx(v, values[samples, ti, sl], REL, LODE_FLOOR, f"values {b}"))
        assert dev.n_gradient_rows[b] == offsets[-1]
        g = dev.gradients(b)
        if g.shape[0] == 0:
            continue
        scale = float(g.abs().max())
        sample_of_row = torch.from_numpy(np.repeat(np.arange(len(samples)), per)).to(g.device)
        resid = torch.zeros((len(samples),) + tuple(g.shape[1:]), dtype=torch.float64, device=g.device)
        resid.index_add_(0, sample_of_row, g.double())
        worst_resid = max(worst_resid, float(resid.abs().max()) / scale)
        got, exp = [], []
        for c, i in enumerate(centers):
            if types[i] != ct:
                continue
            s = int(np.searchsorted(samples, i))
            atoms = np.nonzero((types == nt) | (np.arange(n) == i))[0]
            rows = g[offsets[s]:offsets[s + 1]].double().cpu().numpy().reshape(len(atoms), 3, 2 * l + 1, N)
            for r, j in enumerate(atoms):
                got.append(rows[r])
                exp.append(self_grads[c, ti, :, sl] if j == i else grads[c, j, :, sl])
        if got:
            needed[1] = max(needed[1], assert_block_approx(np.array(got), np.array(exp), REL, LODE_FLOOR,
                                                           f"gradient rows block {b}", scale=scale))
    say(f"  floors needed: values {needed[0]:.1e}, gradient rows {needed[1]:.1e}; translation residual {worst_resid:.1e}")
    assert worst_resid <= 1e-5
