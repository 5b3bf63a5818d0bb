"""GPU tests of the device-resident path (featomic_amd_compute_device, used by the
benchmark): the same keys, labels, values and position gradients as the host
TensorMap of featomic_calculator_compute on the same systems, bit for bit (the
host path runs the same kernels and copies their output)."""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from rascaline_amd.device import DeviceSystems, compute_device
from helpers import random_system

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


SOAP = {
    "cutoff": {"radius": 4.5, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 7}},
}
LODE = {
    "density": {"type": "SmearedPowerLaw", "smearing": 1.0, "exponent": 1},
    "basis": {"type": "TensorProduct", "max_angular": 3, "radial": {"type": "Gto", "max_radial": 3, "radius": 4.0}},
}


def batch(seed, kinds=("periodic", "triclinic", "cluster")):
    rng = np.random.default_rng(seed)
    out = []
    for k, kind in enumerate(kinds):
        pos, types, cell = random_system(rng, 20 + 7 * k, kind, 7.0 if kind != "cluster" else 9.0, types=(1, 6, 8))
        out.append((pos, types, cell))
    return out


def compare(calc, systems, with_labels=True):
    import torch

    host = calc.compute([ra.System(t, p, c) for p, t, c in systems], gradients=["positions"])
    dev = compute_device(calc, DeviceSystems(systems), ["positions"])
    torch.cuda.synchronize()
    assert dev.n_blocks == len(host.keys.values)
    assert np.array_equal(dev.keys, np.asarray(host.keys.values)), (dev.keys, host.keys.values)
    for b, block in enumerate(host.blocks()):
        values = dev.values(b).double().cpu().numpy()
        assert values.shape == block.values.reshape(len(block.samples), -1).shape
        assert np.array_equal(values, block.values.reshape(values.shape)), f"values of block {b}"
        g_host = block.gradient("positions")
        grads = dev.gradients(b).double().cpu().numpy()
        assert np.array_equal(grads, g_host.values.reshape(grads.shape)), f"gradients of block {b}"
        if with_labels:
            assert np.array_equal(dev.sample_labels(b).cpu().numpy(), np.asarray(block.samples.values))
            assert np.array_equal(dev.gradient_labels(b).cpu().numpy(), np.asarray(g_host.samples.values))
        else:
            assert dev.sample_labels(b) is None and dev.gradient_labels(b) is None
    return dev


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("fast", [False, True])
def test_power_spectrum_device_output_is_the_host_output(precision, fast):
    if fast and precision == 64:
        pytest.skip("the fast form is FP32 only")
    say("power spectrum device output, precision", precision, "fast", fast)
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(precision)
    calc.set_fast_form(fast)
    compare(calc, batch(3))


@pytest.mark.parametrize("precision", [32, 64])
def test_spherical_expansion_device_output_is_the_host_output(precision):
    say("spherical expansion device output, precision", precision)
    calc = ra.SphericalExpansion(**SOAP)
    calc.set_precision(precision)
    compare(calc, batch(4))


@pytest.mark.parametrize("precision", [32, 64])
def test_lode_device_output_is_the_host_output(precision):
    say("LODE device output, precision", precision)
    calc = ra.LodeSphericalExpansion(**LODE)
    calc.set_precision(precision)
    compare(calc, batch(5, kinds=("periodic", "triclinic")), with_labels=False)


def test_device_output_against_oracle():
    """The device-resident FP64 power spectrum against the FP64 oracle directly."""
    import torch

    systems = batch(6)
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(64)
    dev = compute_device(calc, DeviceSystems(systems), ["positions"])
    torch.cuda.synchronize()
    ref = orc.soap_power_spectrum([orc.System(p, t, c) for p, t, c in systems], SOAP, ["positions"])
    assert dev.n_blocks == len(ref["blocks"])
    vscale = max(np.abs(b["values"]).max() for b in ref["blocks"])
    gscale = max(np.abs(b["gradients"]["positions"]["values"]).max() for b in ref["blocks"] if
                 b["gradients"]["positions"]["values"].size)
    for b, rb in enumerate(ref["blocks"]):
        v = dev.values(b).cpu().numpy()
        assert np.all(np.abs(v - rb["values"].reshape(v.shape)) <= 1e-10 * (np.abs(rb["values"].reshape(v.shape)) + vscale))
        g = dev.gradients(b).cpu().numpy()
        rg = rb["gradients"]["positions"]["values"].reshape(g.shape)
        assert np.all(np.abs(g - rg) <= 1e-10 * (np.abs(rg) + gscale))


def block_arrays(dev):
    return [(dev.values(b).double().cpu().numpy(), dev.gradients(b).double().cpu().numpy()) for b in range(dev.n_blocks)]


def max_block_error(a, b):
    worst = 0.0
    for (va, ga), (vb, gb) in zip(a, b):
        for x, y in ((va, vb), (ga, gb)):
            if y.size and np.abs(y).max() > 0:
                worst = max(worst, float(np.abs(x - y).max() / np.abs(y).max()))
    return worst


@pytest.mark.parametrize("precision,fast", [(32, False), (32, True), (64, False)])
def test_every_tuned_launch_shape_and_autotune(precision, fast):
    """Every launch shape the autotuner can choose (each window shape of the tiled
    gradient kernel, the register-blocked kernel, 4 to 64 expansion entries per
    chunk) gives the built-in result to rounding (block-relative 1e-5 in FP32, 1e-12
    in FP64, 1e-4 for the fast form, whose shapes beyond its mixed kernel compute the
    plain contraction); the autotuner keeps one of its candidates."""
    import torch

    rng = np.random.default_rng(17)
    systems = []
    for kind, n, side in [("periodic", 60, 8.0), ("triclinic", 45, 7.5), ("cluster", 30, 8.0)]:
        systems.append(random_system(rng, n, kind, side, types=(1, 6, 8)))
    ds = DeviceSystems(systems)
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(precision)
    calc.set_fast_form(fast)
    ref = block_arrays(compute_device(calc, ds, ["positions"]))
    # FP32 rounding; the fast form's own error where a shape falls back to the plain contraction
    bar = 1e-12 if precision == 64 else (1e-4 if fast else 1e-5)
    report = calc.autotune(ds, repeats=1)
    say("autotune report:", report)
    windows = [c["choice"] for c in report["gradient_window"]]
    assert windows[0] == -1 and len(windows) >= 3
    # the fast form is tuned among the shapes that keep its mixed-precision contraction
    assert (windows[-1] == -2) == (not fast)
    assert calc.tuning["gradient_window"] in windows
    assert calc.tuning["expansion_entries"] in [c["entries"] for c in report["expansion_entries"]]
    tuned = block_arrays(compute_device(calc, ds, ["positions"]))
    assert max_block_error(tuned, ref) <= bar
    for w in windows + ([] if not fast else [-2]):
        if any(c["choice"] == w and "error" in c for c in report["gradient_window"]):
            continue
        calc.tuning = {"expansion_entries": 0, "gradient_window": w}
        err = max_block_error(block_arrays(compute_device(calc, ds, ["positions"])), ref)
        say(f"  gradient window {w}: {err:.1e}")
        assert err <= bar, w
    # negative counts: the matrix-core expansion kernel (FP32; FP64 takes the vector-ALU one)
    for e in (4, 16, 32, 64, -8, -16):
        calc.tuning = {"expansion_entries": e, "gradient_window": -1}
        err = max_block_error(block_arrays(compute_device(calc, ds, ["positions"])), ref)
        say(f"  expansion entries {e}: {err:.1e}")
        assert err <= bar, e
    torch.cuda.synchronize()


def test_calculators_on_concurrent_streams():
    """Chunks alternating between two calculators on their own streams (as the benchmark
    pipelines them) give the results each chunk has alone."""
    import torch

    chunks = [DeviceSystems(batch(20 + k)) for k in range(4)]
    calcs = []
    for _ in range(2):
        c = ra.SoapPowerSpectrum(**SOAP)
        c.set_precision(32)
        calcs.append(c)
    alone = []
    for ds in chunks:
        alone.append(block_arrays(compute_device(calcs[0], ds, ["positions"])))
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    for rep in range(2):
        results = [None] * len(chunks)
        # chunks 0 and 1 in flight together, then 2 and 3
        for pair in ((0, 1), (2, 3)):
            devs = [compute_device(calcs[k % 2], chunks[k], ["positions"], stream=streams[k % 2]) for k in pair]
            torch.cuda.synchronize()
            for k, dev in zip(pair, devs):
                results[k] = block_arrays(dev)
        for k in range(len(chunks)):
            assert max_block_error(results[k], alone[k]) == 0.0, (rep, k)
