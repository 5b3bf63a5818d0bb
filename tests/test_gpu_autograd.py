"""GPU tests of the autograd path (rascaline_amd.autograd.soap_values): the backward
is the vector-Jacobian product of the position gradients.

* torch.autograd.gradcheck (fast mode: a random direction, central finite
  differences of the forward in FP64) on small periodic and cluster systems, for
  the power spectrum, the spherical expansion, the radial spectrum and LODE
  (periodic only);
* loss.backward() on Σ W · values against featomic-torch's contraction of the FP64
  oracle's gradient rows with W (1e-10);
* the double backward (featomic-torch's PositionsGrad::backward): gradcheck of the
  backward as a function of dA/dX, and a loss on forces whose gradient with respect
  to the weights is checked against the oracle's gradient rows (FP64 1e-10, FP32
  2e-5 of the largest element).
"""

import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from rascaline_amd.autograd import soap_values
from helpers import random_system

pytestmark = pytest.mark.gpu

SOAP = {
    "cutoff": {"radius": 4.0, "smoothing": {"type": "ShiftedCosine", "width": 1.0}},
    "density": {"type": "Gaussian", "width": 0.4},
    "basis": {"type": "TensorProduct", "max_angular": 3, "radial": {"type": "Gto", "max_radial": 3}},
}


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def small_systems(seed):
    rng = np.random.default_rng(seed)
    a = random_system(rng, 7, "periodic", 5.5, types=(1, 6))
    b = random_system(rng, 5, "cluster", 4.0, types=(1, 8))
    return [a, b]


LODE = {
    "density": {"type": "SmearedPowerLaw", "smearing": 1.0, "exponent": 1},
    "basis": {"type": "TensorProduct", "max_angular": 2, "radial": {"type": "Gto", "max_radial": 2, "radius": 3.0}},
}


@pytest.mark.parametrize("kind", ["power_spectrum", "spherical_expansion", "radial_spectrum", "lode"])
def test_gradcheck(kind):
    import torch

    say("gradcheck", kind)
    systems = small_systems(3)
    if kind == "lode":
        rng = np.random.default_rng(4)
        systems = [random_system(rng, 6, "triclinic", 5.0, types=(1, 8)), random_system(rng, 5, "periodic", 4.5, types=(1, 6))]
        calc = ra.LodeSphericalExpansion(**LODE)
    elif kind == "radial_spectrum":
        calc = ra.SoapRadialSpectrum(cutoff=SOAP["cutoff"], density=SOAP["density"],
                                     basis={"radial": SOAP["basis"]["radial"]})
    else:
        calc = ra.SoapPowerSpectrum(**SOAP) if kind == "power_spectrum" else ra.SphericalExpansion(**SOAP)
    calc.set_precision(64)
    pos = torch.tensor(np.concatenate([s[0] for s in systems]), dtype=torch.float64, device="cuda",
                       requires_grad=True)
    assert torch.autograd.gradcheck(lambda p: soap_values(calc, systems, p)[0], (pos,), eps=1e-6, atol=1e-8,
                                    rtol=1e-5, fast_mode=True)


def test_backward_vs_oracle():
    import torch

    rng = np.random.default_rng(11)
    systems = [random_system(rng, 16, "triclinic", 6.5, types=(1, 6, 8)),
               random_system(rng, 12, "cluster", 7.0, types=(1, 6, 8))]
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(64)
    pos = torch.tensor(np.concatenate([s[0] for s in systems]), dtype=torch.float64, device="cuda",
                       requires_grad=True)
    values, layout = soap_values(calc, systems, pos)
    ref = orc.soap_power_spectrum([orc.System(p, t, c) for p, t, c in systems], SOAP, ["positions"])
    assert np.array_equal(layout.keys, np.asarray(ref["keys"]))
    weights = [rng.normal(size=(layout.n_samples[b], layout.properties[b])) for b in range(len(layout.keys))]
    loss = sum((layout.block(values, b) * torch.as_tensor(weights[b], device="cuda")).sum()
               for b in range(len(layout.keys)))
    loss.backward()
    offsets = np.concatenate([[0], np.cumsum([len(s[1]) for s in systems])])
    expected = np.zeros((offsets[-1], 3))
    for blk, w in zip(ref["blocks"], weights):
        g = blk["gradients"]["positions"]
        rows = g["values"].reshape(len(g["samples"]), 3, -1)
        for r, (sample, system, atom) in enumerate(g["samples"]):
            expected[offsets[system] + atom] += rows[r] @ w[sample]
    got = pos.grad.cpu().numpy()
    scale = np.abs(expected).max()
    say(f"backward vs oracle: {np.abs(got - expected).max() / scale:.2e}")
    assert np.all(np.abs(got - expected) <= 1e-10 * (np.abs(expected) + scale))
    # the forward values themselves
    for b, blk in enumerate(ref["blocks"]):
        v = layout.block(values, b).detach().cpu().numpy()
        assert np.all(np.abs(v - blk["values"]) <= 1e-10 * (np.abs(blk["values"]) + np.abs(blk["values"]).max()))


def force_loss_systems():
    rng = np.random.default_rng(23)
    return [random_system(rng, 12, "triclinic", 6.0, types=(1, 6, 8)),
            random_system(rng, 9, "cluster", 6.0, types=(1, 8))]


@pytest.mark.parametrize("kind", ["power_spectrum", "spherical_expansion", "lode"])
def test_double_backward_gradcheck(kind):
    """The backward as a function of dA/dX, w -> (dX/dr)^T w, is differentiable: its
    gradient is dX/dr . v (featomic-torch's PositionsGrad::backward); gradcheck with
    central finite differences of the (linear) backward in FP64."""
    import torch

    say("double backward gradcheck", kind)
    systems = small_systems(5)
    if kind == "lode":
        rng = np.random.default_rng(6)
        systems = [random_system(rng, 6, "triclinic", 5.0, types=(1, 8)), random_system(rng, 5, "periodic", 4.5, types=(1, 6))]
        calc = ra.LodeSphericalExpansion(**LODE)
    else:
        calc = ra.SoapPowerSpectrum(**SOAP) if kind == "power_spectrum" else ra.SphericalExpansion(**SOAP)
    calc.set_precision(64)
    pos = torch.tensor(np.concatenate([s[0] for s in systems]), dtype=torch.float64, device="cuda",
                       requires_grad=True)
    values, _ = soap_values(calc, systems, pos)
    w = torch.randn(values.shape, dtype=torch.float64, device="cuda", requires_grad=True)

    def backward_of(w):
        return torch.autograd.grad(values, pos, w, create_graph=True)[0]

    assert torch.autograd.gradcheck(backward_of, (w,), eps=1e-6, atol=1e-8, rtol=1e-6, fast_mode=True)


@pytest.mark.parametrize("precision", [64, 32])
def test_force_loss_vs_oracle(precision):
    """Training on forces: E = Σ W X, F = -dE/dr (create_graph), loss = Σ u . F; then
    dloss/dW = -(dX/dr) . u per sample, which the FP64 oracle's gradient rows give."""
    import torch

    systems = force_loss_systems()
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(precision)
    pos = torch.tensor(np.concatenate([s[0] for s in systems]), dtype=torch.float64, device="cuda",
                       requires_grad=True)
    values, layout = soap_values(calc, systems, pos)
    rng = np.random.default_rng(3)
    weights = torch.tensor(rng.normal(size=values.shape), dtype=values.dtype, device="cuda", requires_grad=True)
    energy = (values * weights).sum()
    forces = -torch.autograd.grad(energy, pos, create_graph=True)[0]
    u = torch.tensor(rng.normal(size=forces.shape), dtype=torch.float64, device="cuda")
    loss = (forces * u).sum()
    import rascaline_amd.autograd as ag

    ag._WARNED[0] = False  # the warning is given once per process
    with pytest.warns(UserWarning, match="second derivatives with respect to positions"):
        loss.backward()
    ref = orc.soap_power_spectrum([orc.System(p, t, c) for p, t, c in systems], SOAP, ["positions"])
    assert np.array_equal(layout.keys, np.asarray(ref["keys"]))
    offsets = np.concatenate([[0], np.cumsum([len(s[1]) for s in systems])])
    un = u.cpu().numpy()
    worst = 0.0
    for b, blk in enumerate(ref["blocks"]):
        g = blk["gradients"]["positions"]
        rows = g["values"].reshape(len(g["samples"]), 3, -1)
        expected = np.zeros((len(blk["samples"]), rows.shape[-1]))
        for r, (sample, system, atom) in enumerate(g["samples"]):
            expected[sample] -= un[offsets[system] + atom] @ rows[r]
        got = layout.block(weights.grad, b).double().cpu().numpy()
        scale = max(np.abs(expected).max(), 1e-300)
        bar = 1e-10 if precision == 64 else 2e-5
        assert np.all(np.abs(got - expected) <= bar * (np.abs(expected) + scale)), b
        worst = max(worst, float(np.abs(got - expected).max() / scale))
    say(f"force-loss weight gradient vs oracle rows, FP{precision}: {worst:.1e}")


def periodic_systems(seed):
    rng = np.random.default_rng(seed)
    return [random_system(rng, 7, "triclinic", 5.5, types=(1, 6)), random_system(rng, 6, "periodic", 5.0, types=(1, 8))]


def cell_tensor(systems, requires_grad=True):
    import torch

    return torch.tensor(np.stack([np.asarray(s[2], dtype=np.float64) for s in systems]), dtype=torch.float64,
                        device="cuda", requires_grad=requires_grad)


@pytest.mark.parametrize("kind", ["power_spectrum", "spherical_expansion"])
def test_cell_gradcheck(kind):
    """featomic-torch's CellGrad: values differentiable with respect to the cells at
    fixed Cartesian positions; gradcheck against central finite differences of the
    forward (FP64), with and without the positions."""
    import torch

    say("cell gradcheck", kind)
    systems = periodic_systems(21)
    calc = ra.SoapPowerSpectrum(**SOAP) if kind == "power_spectrum" else ra.SphericalExpansion(**SOAP)
    calc.set_precision(64)
    pos = torch.tensor(np.concatenate([s[0] for s in systems]), dtype=torch.float64, device="cuda")
    cell = cell_tensor(systems)
    values, _ = soap_values(calc, systems, pos, cell)
    w = torch.tensor(np.random.default_rng(4).normal(size=values.shape), dtype=torch.float64, device="cuda")

    def energy_of_cell(c):
        v, _ = soap_values(calc, systems, pos, c)
        return (v * w).sum()

    assert torch.autograd.gradcheck(energy_of_cell, (cell,), eps=1e-6, atol=1e-7, rtol=1e-5)

    pos_g = pos.clone().requires_grad_(True)

    def energy_of_both(p, c):
        v, _ = soap_values(calc, systems, p, c)
        return (v * w).sum()

    assert torch.autograd.gradcheck(energy_of_both, (pos_g, cell), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_cell_backward_vs_oracle():
    """dA/dH per system = Σ over its samples of W · dX/dH with the FP64 oracle's cell
    gradients (CellGrad::forward, autograd.cpp:428-489)."""
    import torch

    systems = periodic_systems(22)
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(64)
    pos = torch.tensor(np.concatenate([s[0] for s in systems]), dtype=torch.float64, device="cuda")
    cell = cell_tensor(systems)
    values, layout = soap_values(calc, systems, pos, cell)
    rng = np.random.default_rng(8)
    w = torch.tensor(rng.normal(size=values.shape), dtype=torch.float64, device="cuda")
    (values * w).sum().backward()
    ref = orc.soap_power_spectrum([orc.System(p, t, c) for p, t, c in systems], SOAP, ["cell"])
    assert np.array_equal(layout.keys, np.asarray(ref["keys"]))
    expected = np.zeros((len(systems), 3, 3))
    for b, blk in enumerate(ref["blocks"]):
        g = blk["gradients"]["cell"]["values"].reshape(len(blk["samples"]), 3, 3, -1)
        wb = layout.block(w, b).cpu().numpy()
        for s, (system, _) in enumerate(blk["samples"]):
            expected[system] += np.einsum("abp,p->ab", g[s], wb[s])
    got = cell.grad.cpu().numpy()
    scale = np.abs(expected).max()
    assert np.all(np.abs(got - expected) <= 1e-10 * (np.abs(expected) + scale)), (got, expected)


@pytest.mark.parametrize("precision", [64, 32])
def test_cell_double_backward(precision):
    """The cell backward as a function of dA/dX is differentiable (CellGrad::backward,
    autograd.cpp:493-560): a loss on the cell gradient (a stress-like quantity) reaches
    the weights; gradcheck of w -> dA/dH(w) in FP64, and the second derivative with
    respect to the cell warns once."""
    import torch

    import rascaline_amd.autograd as ag

    systems = periodic_systems(23)
    calc = ra.SoapPowerSpectrum(**SOAP)
    calc.set_precision(precision)
    pos = torch.tensor(np.concatenate([s[0] for s in systems]), dtype=torch.float64, device="cuda")
    cell = cell_tensor(systems)
    values, _ = soap_values(calc, systems, pos, cell)
    if precision == 64:
        w = torch.randn(values.shape, dtype=torch.float64, device="cuda", requires_grad=True)

        def backward_of(w):
            return torch.autograd.grad(values, cell, w, create_graph=True)[0]

        assert torch.autograd.gradcheck(backward_of, (w,), eps=1e-6, atol=1e-8, rtol=1e-6)
    weights = torch.randn(values.shape, dtype=values.dtype, device="cuda", requires_grad=True)
    stress = torch.autograd.grad((values * weights).sum(), cell, create_graph=True)[0]
    ag._WARNED_CELL[0] = False
    with pytest.warns(UserWarning, match="second derivatives with respect to cell"):
        (stress ** 2).sum().backward()
    assert weights.grad is not None and float(weights.grad.abs().max()) > 0.0
