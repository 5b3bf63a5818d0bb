"""GPU checks at the full sizes of the two single-cell workloads.

* Bulk power spectrum: one periodic cubic cell of 20 000 atoms at 0.1 atoms/A^3,
  5 types uniform, rc 6 A (ShiftedCosine 0.5 A), Gaussian width 0.3 A, 12 radial
  functions, l <= 8, FP32 plain and fast (bfloat16 hi + lo MFMA) forms.
  - values of the whole cell against the FP64 oracle: error of every
    (l, center type, neighbour type pair) slice, max|a - b| / max|b| <= 1e-4;
  - position gradients (about 9 M rows, 145 GB, kept in device memory):
    translation invariance, Σ_j dPS_i/dr_j = 0 over each sample's rows, within
    1e-4 of the slice's largest gradient; and the fast form against the plain one
    on the first rows of every block, within 1e-4 per slice.
* Spherical expansion through the learned radial integral: one cell of 100 000
  atoms at 0.1 atoms/A^3 (L = 100 A), 4 types, rc 6 A, 12 radial functions,
  l <= 8, FP32, network trained on 1e6 distances and accepted by the spline judge.
  - position gradients (about 9.4 M rows, 110 GB) translation invariant per
    (l, block) within 1e-4; values and the first rows of every gradient block
    within 1e-4 of the spline path per (l, block);
  - values on a 10 000-atom cell of the same density against the FP64 oracle,
    within 1e-4 per (l, block).
"""

import gc
import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from rascaline_amd.datasets import bulk_cell
from rascaline_amd.device import DeviceSystems, compute_device

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


BULK = {
    "cutoff": {"radius": 6.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 8, "radial": {"type": "Gto", "max_radial": 11}},
}
L, N = 9, 12


def slices_ps(l):
    return slice(l * N * N, (l + 1) * N * N)


def slice_errors(actual, expected, slices):
    """{l: max|a - b| / max|b|} over the property slices of one block."""
    out = {}
    for l, sl in enumerate(slices):
        scale = float(np.abs(expected[..., sl]).max()) if expected.size else 0.0
        if scale > 0:
            out[l] = float(np.abs(actual[..., sl] - expected[..., sl]).max()) / scale
    return out


def translation_residual(dev, b, slices):
    """{l: max over samples of |Σ_rows g| / max|g|} for block b of a device result."""
    import torch

    g = dev.gradients(b)
    sample = dev.gradient_labels(b)[:, 0].long()
    total = torch.zeros((dev.n_samples[b],) + tuple(g.shape[1:]), dtype=torch.float64, device=g.device)
    step = 65536
    for k in range(0, g.shape[0], step):
        total.index_add_(0, sample[k:k + step], g[k:k + step].double())
    out = {}
    for l, sl in enumerate(slices):
        scale = float(g[..., sl].abs().max())
        if scale > 0:
            out[l] = float(total[..., sl].abs().max()) / scale
    return out


def first_rows(dev, b, count):
    return dev.gradients(b)[:count].clone()


@pytest.fixture(scope="module")
def bulk():
    return bulk_cell(seed=11, n_atoms=20000, density=0.1, types=(1, 6, 7, 8, 9))


@pytest.mark.timeout(900)
def test_bulk_values_full_size_vs_oracle(bulk):
    (pos, types, cell), = bulk
    say("bulk values, 20000 atoms: oracle")
    ref = orc.soap_power_spectrum([orc.System(pos, types, cell)], BULK, [])
    slices = [slices_ps(l) for l in range(L)]
    for fast in (False, True):
        calc = ra.SoapPowerSpectrum(**BULK)
        calc.set_precision(32)
        calc.set_fast_form(fast)
        tm = calc.compute([ra.System(types, pos, cell)])
        assert np.array_equal(np.asarray(tm.keys.values), np.asarray(ref["keys"]))
        worst = 0.0
        for block, rb in zip(tm.blocks(), ref["blocks"]):
            errs = slice_errors(block.values, rb["values"], slices)
            assert len(errs) == L
            worst = max(worst, max(errs.values()))
        say("  fast" if fast else "  plain", f"worst (l, block) error of the values {worst:.2e}")
        assert worst <= 1e-4


@pytest.mark.timeout(900)
def test_bulk_gradients_full_size(bulk):
    import torch

    slices = [slices_ps(l) for l in range(L)]
    keep = 20000
    kept = {}
    for fast in (False, True):
        calc = ra.SoapPowerSpectrum(**BULK)
        calc.set_precision(32)
        calc.set_fast_form(fast)
        dev = compute_device(calc, DeviceSystems(bulk), ["positions"])
        torch.cuda.synchronize()
        say("  fast" if fast else "  plain", f"{dev.total_rows} gradient rows in {dev.n_blocks} blocks")
        assert dev.total_rows > 5_000_000
        worst_t, worst_f = 0.0, 0.0
        for b in range(dev.n_blocks):
            res = translation_residual(dev, b, slices)
            worst_t = max(worst_t, max(res.values()))
            rows = first_rows(dev, b, keep)
            if not fast:
                kept[b] = rows
            else:
                errs = slice_errors(rows.cpu().numpy(), kept[b].cpu().numpy(), slices)
                worst_f = max(worst_f, max(errs.values()))
        say("    translation residual (worst (l, block))", f"{worst_t:.2e}",
            *(["fast vs plain on the first rows", f"{worst_f:.2e}"] if fast else []))
        assert worst_t <= 1e-4
        assert worst_f <= 1e-4
        del dev, calc
        gc.collect()
        torch.cuda.synchronize()


LARGE = BULK


@pytest.mark.timeout(900)
def test_surrogate_large_cell_full_size():
    import torch

    cells = bulk_cell(seed=12, n_atoms=100000, density=0.1, types=(1, 6, 7, 8))
    calc = ra.SphericalExpansion(**LARGE)
    calc.set_precision(32)
    report = calc.train_radial_surrogate()
    say(f"  network accepted {report['accepted']}: value error {report['max_value_error']:.2e}, derivative "
        f"error {report['max_derivative_error']:.2e}, {report['train_ms']:.0f} ms")
    assert report["accepted"]
    systems = DeviceSystems(cells)

    def slices_se(b, dev):
        # SE blocks are one l each: the whole property range is the slice
        return [slice(0, dev.properties[b])]

    dev = compute_device(calc, systems, ["positions"])
    torch.cuda.synchronize()
    say(f"  surrogate: {dev.total_rows} gradient rows in {dev.n_blocks} blocks")
    assert dev.total_rows > 5_000_000
    worst_t = 0.0
    kept_v, kept_g = {}, {}
    for b in range(dev.n_blocks):
        res = translation_residual(dev, b, slices_se(b, dev))
        worst_t = max(worst_t, max(res.values()))
        kept_v[b] = dev.values(b).clone()
        kept_g[b] = first_rows(dev, b, 20000)
    say(f"    translation residual (worst block) {worst_t:.2e}")
    assert worst_t <= 1e-4

    calc.use_radial_surrogate(False)
    dev = compute_device(calc, systems, ["positions"])
    torch.cuda.synchronize()
    worst_v, worst_g = 0.0, 0.0
    for b in range(dev.n_blocks):
        sl = slices_se(b, dev)
        ev = slice_errors(kept_v[b].cpu().numpy(), dev.values(b).cpu().numpy(), sl)
        eg = slice_errors(kept_g[b].cpu().numpy(), first_rows(dev, b, 20000).cpu().numpy(), sl)
        worst_v = max([worst_v] + list(ev.values()))
        worst_g = max([worst_g] + list(eg.values()))
    say(f"    surrogate vs spline (worst block): values {worst_v:.2e}, first gradient rows {worst_g:.2e}")
    assert worst_v <= 1e-4 and worst_g <= 1e-4
    calc.use_radial_surrogate(True)
    del dev
    gc.collect()

    # the network against the FP64 oracle on a 10 000-atom cell of the same density
    (pos, types, cell), = bulk_cell(seed=13, n_atoms=10000, density=0.1, types=(1, 6, 7, 8))
    ref = orc.soap_spherical_expansion([orc.System(pos, types, cell)], LARGE, [])
    tm = calc.compute([ra.System(types, pos, cell)])
    worst = 0.0
    for block, rb in zip(tm.blocks(), ref["blocks"]):
        a = block.values.reshape(len(block.samples), -1)
        e = rb["values"].reshape(a.shape)
        worst = max(worst, float(np.abs(a - e).max()) / float(np.abs(e).max()))
    say(f"    10 000-atom cell, network vs FP64 oracle (worst block): values {worst:.2e}")
    assert worst <= 1e-4
