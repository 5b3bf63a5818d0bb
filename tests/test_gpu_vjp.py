"""GPU tests of the vector-Jacobian product of the position gradients
(featomic_amd_positions_vjp): dA/dr = Σ_blocks Σ_rows (dX/dr)[row] · dA/dX[sample of
row], the contraction of featomic-torch's PositionsGrad::forward
(featomic-torch/src/autograd.cpp:237-300), computed on the GPU without the gradient
rows.

* against the same contraction of the FP64 oracle's gradient rows (the reference's
  formula, on the host): FP64 within 1e-10 (|a - b| <= 1e-10 (|b| + max|b|)), FP32
  within 1e-4 max|b|, power spectrum and spherical expansion, periodic, triclinic
  and cluster systems in one batch;
* against the library's own materialised gradient rows contracted with torch
  (FP64, 1e-12).
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from rascaline_amd.device import DeviceSystems, compute_device, positions_vjp
from helpers import random_system

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


SOAP = {
    "cutoff": {"radius": 4.5, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5}},
}


def batch(seed):
    rng = np.random.default_rng(seed)
    out = []
    for k, kind in enumerate(("periodic", "triclinic", "cluster")):
        pos, types, cell = random_system(rng, 18 + 6 * k, kind, 7.0 if kind != "cluster" else 8.0, types=(1, 6, 8))
        out.append((pos, types, cell))
    return out


def reference_vjp(ref, weights, offsets):
    """featomic-torch's PositionsGrad::forward on the oracle's gradient rows."""
    n_atoms = offsets[-1]
    out = np.zeros((n_atoms, 3))
    for blk, w in zip(ref["blocks"], weights):
        g = blk["gradients"]["positions"]
        rows = g["values"].reshape(len(g["samples"]), 3, -1)
        for r, (sample, system, atom) in enumerate(g["samples"]):
            out[offsets[system] + atom] += rows[r] @ w[sample]
    return out


def device_weights(dev, weights, torch, dtype):
    """dA/dX laid out like the device values: block b at its value offset."""
    flat = torch.zeros(dev.value_size, dtype=dtype, device="cuda")
    for b in range(dev.n_blocks):
        off, size = dev.value_offset(b), dev.n_samples[b] * dev.properties[b]
        flat[off:off + size] = torch.as_tensor(weights[b].reshape(-1), dtype=dtype, device="cuda")
    return flat


@pytest.mark.parametrize("kind", ["power_spectrum", "spherical_expansion"])
@pytest.mark.parametrize("precision", [64, 32])
def test_positions_vjp_vs_oracle(kind, precision):
    import torch

    say("positions VJP", kind, "precision", precision)
    systems = batch(41 if kind == "power_spectrum" else 42)
    osys = [orc.System(p, t, c) for p, t, c in systems]
    if kind == "power_spectrum":
        ref = orc.soap_power_spectrum(osys, SOAP, ["positions"])
        calc = ra.SoapPowerSpectrum(**SOAP)
    else:
        ref = orc.soap_spherical_expansion(osys, SOAP, ["positions"])
        calc = ra.SphericalExpansion(**SOAP)
    calc.set_precision(precision)
    rng = np.random.default_rng(5)
    weights = [rng.normal(size=(len(b["samples"]), int(np.prod(b["values"].shape[1:])))) for b in ref["blocks"]]
    offsets = np.concatenate([[0], np.cumsum([len(t) for _, t, _ in systems])])
    expected = reference_vjp(ref, weights, offsets)

    ds = DeviceSystems(systems)
    dev = compute_device(calc, ds)
    assert dev.n_blocks == len(ref["blocks"])
    assert np.array_equal(dev.keys, np.asarray(ref["keys"]))
    dtype = torch.float64 if precision == 64 else torch.float32
    w = device_weights(dev, weights, torch, dtype)
    got = positions_vjp(calc, ds, w).double().cpu().numpy()
    scale = np.abs(expected).max()
    err = np.abs(got - expected).max() / scale
    say(f"  max |a - b| / max|b| = {err:.2e}")
    if precision == 64:
        assert np.all(np.abs(got - expected) <= 1e-10 * (np.abs(expected) + scale))
    else:
        assert err <= 1e-4


def test_positions_vjp_matches_materialised_rows():
    import torch

    systems = batch(43)
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(64)
    ds = DeviceSystems(systems)
    dev = compute_device(calc, ds, ["positions"])
    torch.cuda.synchronize()
    rng = np.random.default_rng(9)
    weights = [rng.normal(size=(dev.n_samples[b], dev.properties[b])) for b in range(dev.n_blocks)]
    offsets = torch.as_tensor(np.concatenate([[0], np.cumsum([len(t) for _, t, _ in systems])]), device="cuda")
    expected = torch.zeros((ds.n_atoms, 3), dtype=torch.float64, device="cuda")
    for b in range(dev.n_blocks):
        g = dev.gradients(b)
        labels = dev.gradient_labels(b).long()
        wb = torch.as_tensor(weights[b], device="cuda")
        contrib = torch.einsum("rxp,rp->rx", g, wb[labels[:, 0]])
        expected.index_add_(0, offsets[labels[:, 1]] + labels[:, 2], contrib)
    expected = expected.cpu().numpy()
    w = device_weights(dev, weights, torch, torch.float64)
    got = positions_vjp(calc, ds, w).cpu().numpy()
    scale = np.abs(expected).max()
    assert np.all(np.abs(got - expected) <= 1e-12 * (np.abs(expected) + scale))


LODE = {
    "density": {"type": "SmearedPowerLaw", "smearing": 1.0, "exponent": 1},
    "basis": {"type": "TensorProduct", "max_angular": 3, "radial": {"type": "Gto", "max_radial": 3, "radius": 4.0}},
}


@pytest.mark.parametrize("precision", [64, 32])
def test_lode_positions_vjp_vs_oracle(precision):
    """LODE: the VJP from the structure factors against featomic-torch's contraction of
    the oracle's N^2 gradient rows (two periodic systems, three types)."""
    import torch

    say("LODE positions VJP, precision", precision)
    rng = np.random.default_rng(44)
    systems = [random_system(rng, 14, "triclinic", 6.5, types=(1, 6, 8)),
               random_system(rng, 10, "periodic", 6.0, types=(1, 8))]
    ref = orc.lode_spherical_expansion([orc.System(p, t, c) for p, t, c in systems], LODE, ["positions"])
    calc = ra.LodeSphericalExpansion(**LODE)
    calc.set_precision(precision)
    weights = [rng.normal(size=(len(b["samples"]), int(np.prod(b["values"].shape[1:])))) for b in ref["blocks"]]
    offsets = np.concatenate([[0], np.cumsum([len(t) for _, t, _ in systems])])
    expected = reference_vjp(ref, weights, offsets)
    ds = DeviceSystems(systems)
    dev = compute_device(calc, ds)
    assert dev.n_blocks == len(ref["blocks"])
    assert np.array_equal(dev.keys, np.asarray(ref["keys"]))
    dtype = torch.float64 if precision == 64 else torch.float32
    got = positions_vjp(calc, ds, device_weights(dev, weights, torch, dtype)).double().cpu().numpy()
    scale = np.abs(expected).max()
    err = np.abs(got - expected).max() / scale
    say(f"  max |a - b| / max|b| = {err:.2e}")
    if precision == 64:
        assert np.all(np.abs(got - expected) <= 1e-10 * (np.abs(expected) + scale))
    else:
        assert err <= 1e-4


@pytest.mark.parametrize("precision", [64, 32])
def test_radial_spectrum_positions_vjp_vs_oracle(precision):
    """The radial spectrum (the l = 0 expansion, relabelled) through the same VJP entry
    point, against featomic-torch's contraction of the oracle's gradient rows."""
    import torch

    say("radial spectrum positions VJP, precision", precision)
    hypers = {"cutoff": SOAP["cutoff"], "density": SOAP["density"],
              "basis": {"radial": {"type": "Gto", "max_radial": 5}}}
    systems = batch(45)
    ref = orc.soap_radial_spectrum([orc.System(p, t, c) for p, t, c in systems], hypers, ["positions"])
    calc = ra.SoapRadialSpectrum(**hypers)
    calc.set_precision(precision)
    rng = np.random.default_rng(6)
    weights = [rng.normal(size=(len(b["samples"]), int(np.prod(b["values"].shape[1:])))) for b in ref["blocks"]]
    offsets = np.concatenate([[0], np.cumsum([len(t) for _, t, _ in systems])])
    expected = reference_vjp(ref, weights, offsets)
    ds = DeviceSystems(systems)
    dev = compute_device(calc, ds)
    assert dev.n_blocks == len(ref["blocks"])
    assert np.array_equal(dev.keys, np.asarray(ref["keys"]))
    dtype = torch.float64 if precision == 64 else torch.float32
    got = positions_vjp(calc, ds, device_weights(dev, weights, torch, dtype)).double().cpu().numpy()
    scale = np.abs(expected).max()
    err = np.abs(got - expected).max() / scale
    say(f"  max |a - b| / max|b| = {err:.2e}")
    if precision == 64:
        assert np.all(np.abs(got - expected) <= 1e-10 * (np.abs(expected) + scale))
    else:
        assert err <= 1e-4
