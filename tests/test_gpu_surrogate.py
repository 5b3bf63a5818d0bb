"""GPU tests of the learned radial integral.

* The device training kernels (FP32 forward + backward) against the FP64 oracle's
  loss and gradient: |g - g_ref| <= 1e-4 max|g_ref|, loss within 1e-5 relative.
* The device inference kernel (FP16 hi + lo operands on MFMA, FP32 accumulation)
  against the oracle's FP64 forward: |y - y_ref| <= 1e-4 max|y_ref|.
* Training on the GTO radial integral of the large-cell workload (cutoff 6 A,
  Gaussian width 0.3 A, 12 radial functions, l <= 8): accepted by the spline judge
  with every per-(l, n) error <= 2e-4 (values) and 5e-4 (derivatives) of that
  channel's largest magnitude, and at least 10x better than the same network
  without the Adam phase.
* Spherical expansion with position gradients through the network against the FP64
  oracle (block-relative error <= 1e-4, values and gradients); switching back to
  the spline restores the plain results (FP32 within 1e-4 block-relative, FP64
  within 1e-10 (|b| + max|b|)); the spline fallback when the judge rejects the
  network (bit-identical to a calculator without one).
* What the network gains over the method without it: with spline_accuracy null the
  calculator evaluates the exact GTO integral (FP64 1F1 per entry, l and n); the
  network replaces that evaluation at least 3x faster on a 3000-atom cell, within
  1e-4 of the FP64 oracle, and in the time of the spline.
"""

import ctypes
import sys

import numpy as np
import pytest

import featomic_oracle as orc
import rascaline_amd as ra
from rascaline_amd import _c_lib as C
from helpers import block_relative_error, compare_tensormaps, oracle_to_tensormap, random_system

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def device_loss_gradient(params, x, t, n_out):
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    grad = np.zeros(params.size, dtype=np.float32)
    loss = ctypes.c_double()
    C.check(C.lib().featomic_amd_mlp_loss_gradient(n_out, x.size, params.ctypes.data_as(ctypes.c_void_p),
                                                   x.ctypes.data_as(ctypes.c_void_p),
                                                   t.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss),
                                                   grad.ctypes.data_as(ctypes.c_void_p)))
    return loss.value, grad


def device_forward(params, x, n_out):
    params = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros((x.size, n_out), dtype=np.float32)
    C.check(C.lib().featomic_amd_mlp_forward(n_out, x.size, params.ctypes.data_as(ctypes.c_void_p),
                                             x.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p)))
    return out


def random_params(n_out, seed):
    rng = np.random.default_rng(seed)
    p = 0.3 * rng.normal(size=orc.mlp_n_params(n_out))
    # first layer as in training: transitions spread over [-1, 1]
    H = orc.MLP_HIDDEN
    p[:H] = 8.0 * np.where(np.arange(H) % 2 == 0, 1.0, -1.0)
    p[H:2 * H] = -p[:H] * np.linspace(-1.05, 1.05, H)
    return p.astype(np.float32)


@pytest.mark.parametrize("n_out,n_samples", [(6, 100), (216, 77), (216, 1000), (130, 33)])
def test_training_kernels_vs_oracle(n_out, n_samples):
    say("training kernels, n_out", n_out, "samples", n_samples)
    rng = np.random.default_rng(n_out + n_samples)
    params = random_params(n_out, n_samples)
    x = rng.uniform(-1, 1, n_samples).astype(np.float32)
    t = rng.normal(size=(n_samples, n_out)).astype(np.float32)
    loss, grad = device_loss_gradient(params, x, t, n_out)
    ref_loss, ref_grad = orc.mlp_loss_gradient(params.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                                               n_out)
    assert abs(loss - ref_loss) <= 1e-5 * abs(ref_loss)
    err = np.abs(grad - ref_grad).max() / np.abs(ref_grad).max()
    say("  gradient error", err)
    assert err <= 1e-4


@pytest.mark.parametrize("n_out", [216, 130, 6])
def test_inference_kernel_vs_oracle(n_out):
    say("inference kernel, n_out", n_out)
    rng = np.random.default_rng(7 + n_out)
    params = random_params(n_out, 11 * n_out)
    x = np.concatenate([rng.uniform(-1, 1, 997), [-1.0, 1.0, 0.0]]).astype(np.float32)
    y = device_forward(params, x, n_out)
    ref = orc.mlp_forward(params.astype(np.float64), x.astype(np.float64), n_out)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    say("  forward error", err)
    assert err <= 1e-4


LARGE_CELL = {
    "cutoff": {"radius": 6.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 8, "radial": {"type": "Gto", "max_radial": 11}},
}

_TRAINED = {}


def trained_calculator(precision=32):
    """Spherical expansion on the large-cell hyper-parameters with its network trained
    once per precision (defaults: 1e6 samples, 6000 Adam steps)."""
    if precision not in _TRAINED:
        calc = ra.SphericalExpansion(**LARGE_CELL)
        calc.set_precision(precision)
        report = calc.train_radial_surrogate()
        say(f"  trained ({precision} bit): accepted {report['accepted']}, max value error "
            f"{report['max_value_error']:.2e}, max derivative error {report['max_derivative_error']:.2e}, "
            f"{report['train_ms']:.0f} ms, last Adam loss {report['adam_loss']:.2e}")
        _TRAINED[precision] = (calc, report)
    return _TRAINED[precision]


def test_surrogate_training_accepted_and_gain():
    say("surrogate training on the large-cell radial integral")
    calc, report = trained_calculator()
    assert report["accepted"]
    assert report["value_errors"].shape == (9, 12)
    assert report["max_value_error"] <= 2e-4 and report["max_derivative_error"] <= 5e-4
    # the same network and output-layer solve without the Adam phase
    untrained = ra.SphericalExpansion(**LARGE_CELL)
    base = untrained.train_radial_surrogate(steps=0, tolerance=1.0)
    say(f"  without Adam: value {base['max_value_error']:.2e}, derivative {base['max_derivative_error']:.2e}")
    assert report["max_value_error"] <= 0.1 * base["max_value_error"]
    assert report["max_derivative_error"] <= 0.1 * base["max_derivative_error"]
    params = calc.radial_surrogate_parameters()
    assert params.size == orc.mlp_n_params(216) + 216
    assert np.all(np.isfinite(params))


def large_cell_systems(seed):
    rng = np.random.default_rng(seed)
    out = []
    for n, L in [(40, 8.0), (28, 7.0)]:
        pos, types, cell = random_system(rng, n, "periodic", L, types=(1, 6, 7, 8))
        out.append(orc.System(pos, types, cell))
    return out


@pytest.mark.parametrize("precision", [32, 64])
def test_spherical_expansion_through_surrogate(precision):
    say("spherical expansion through the surrogate, precision", precision)
    calc, report = trained_calculator(precision)
    systems = large_cell_systems(5)
    tm = calc.compute([ra.System(s.types, s.positions, s.cell) for s in systems], gradients=["positions"])
    expected = oracle_to_tensormap(orc.soap_spherical_expansion(systems, LARGE_CELL, ["positions"]))
    ev = block_relative_error(tm, expected)
    eg = block_relative_error(tm, expected, "positions")
    say(f"  block-relative error: values {ev:.2e}, gradients {eg:.2e}")
    assert ev <= 1e-4 and eg <= 1e-4
    # switching back to the spline gives the plain computation
    calc.use_radial_surrogate(False)
    try:
        plain = calc.compute([ra.System(s.types, s.positions, s.cell) for s in systems], gradients=["positions"])
    finally:
        calc.use_radial_surrogate(True)
    if precision == 32:
        assert block_relative_error(plain, expected) <= 1e-4
        assert block_relative_error(plain, expected, "positions") <= 1e-4
    else:
        # the FP64 parity bar of the spline path: |a - b| <= 1e-10 (|b| + max|b|)
        compare_tensormaps(plain, expected, 1e-10, 1e-10, gradients=["positions"])


def test_judge_rejects_and_spline_stays():
    say("judge rejection and spline fallback")
    hypers = {
        "cutoff": {"radius": 5.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
        "density": {"type": "Gaussian", "width": 0.3},
        "basis": {"type": "TensorProduct", "max_angular": 3, "radial": {"type": "Gto", "max_radial": 3}},
    }
    systems = large_cell_systems(9)
    ra_sys = [ra.System(s.types, s.positions, s.cell) for s in systems]
    plain = ra.SphericalExpansion(**hypers)
    plain.set_precision(32)
    ref = plain.compute(ra_sys, gradients=["positions"])
    calc = ra.SphericalExpansion(**hypers)
    calc.set_precision(32)
    report = calc.train_radial_surrogate(n_samples=20000, steps=50, tolerance=1e-9)
    assert not report["accepted"]
    with pytest.raises(ra.FeatomicError, match="no accepted radial surrogate"):
        calc.use_radial_surrogate(True)
    out = calc.compute(ra_sys, gradients=["positions"])
    for ba, bb in zip(out.blocks(), ref.blocks()):
        assert np.array_equal(ba.values, bb.values)
        assert np.array_equal(ba.gradient("positions").values, bb.gradient("positions").values)


def test_gain_over_the_exact_radial_integral():
    """The network in place of the exact radial integral (spline_accuracy null): faster
    by at least 3x on the spherical expansion + gradients of a 3000-atom cell, within
    1e-4 of the FP64 oracle; and in the time of the spline."""
    import copy
    import time

    import torch
    from rascaline_amd.device import DeviceSystems, compute_device

    exact_hypers = copy.deepcopy(LARGE_CELL)
    exact_hypers["basis"]["spline_accuracy"] = None
    rng = np.random.default_rng(1)
    pos, types, cell = random_system(rng, 3000, "periodic", (3000 / 0.1) ** (1 / 3), types=(1, 6, 7, 8), min_dist=0.7)
    batch = DeviceSystems([(pos, types, cell)])

    def best_ms(calc, reps=5):
        compute_device(calc, batch, ["positions"])
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            compute_device(calc, batch, ["positions"])
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return 1e3 * best

    exact = ra.SphericalExpansion(**exact_hypers)
    exact.set_precision(32)
    network = ra.SphericalExpansion(**exact_hypers)
    network.set_precision(32)
    report = network.train_radial_surrogate()
    assert report["accepted"]
    spline, _ = trained_calculator()
    spline.use_radial_surrogate(False)
    try:
        t_exact, t_network, t_spline = best_ms(exact), best_ms(network), best_ms(spline)
    finally:
        spline.use_radial_surrogate(True)
    systems = large_cell_systems(5)
    expected = oracle_to_tensormap(orc.soap_spherical_expansion(systems, LARGE_CELL, ["positions"]))
    tm = network.compute([ra.System(s.types, s.positions, s.cell) for s in systems], gradients=["positions"])
    ev, eg = block_relative_error(tm, expected), block_relative_error(tm, expected, "positions")
    say(f"  3000-atom cell: exact {t_exact:.2f} ms, network {t_network:.2f} ms, spline {t_spline:.2f} ms; "
        f"network error values {ev:.2e}, gradients {eg:.2e}")
    assert t_network * 3.0 <= t_exact
    assert t_network <= 1.25 * t_spline
    assert ev <= 1e-4 and eg <= 1e-4
