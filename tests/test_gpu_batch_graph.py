"""GPU tests of the batch graph (featomic_amd_batch_graph_*): one hipGraph holding
a pass of SOAP calculators over several device batches. Every replay must give,
bit for bit, what featomic_amd_compute_device gives eagerly on the same batches
with the same launch shapes (the graph holds the same kernels). Each test replays
several times, with eager work and synchronisations between the replays, and
clears the output buffers before each replay so that a replay that skipped work
would show."""

import sys

import numpy as np
import pytest

import rascaline_amd as ra
from rascaline_amd.datasets import batched_uneven
from rascaline_amd.device import BatchGraph, DeviceSystems, compute_device

pytestmark = pytest.mark.gpu

PS = {
    "cutoff": {"radius": 5.0, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 6, "radial": {"type": "Gto", "max_radial": 7}},
}
SE = {
    "cutoff": {"radius": 4.5, "smoothing": {"type": "ShiftedCosine", "width": 0.5}},
    "density": {"type": "Gaussian", "width": 0.3},
    "basis": {"type": "TensorProduct", "max_angular": 4, "radial": {"type": "Gto", "max_radial": 5}},
}


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def make_batches(n_batches, per_batch, offset=0):
    systems = batched_uneven(seed=0, n_structures=offset + n_batches * per_batch)[offset:]
    return [DeviceSystems(systems[k * per_batch:(k + 1) * per_batch]) for k in range(n_batches)]


def snapshot(result):
    """Host copies of a DeviceResult: keys, then per block labels, values, gradients."""
    blocks = []
    for b in range(result.n_blocks):
        blocks.append((result.sample_labels(b).cpu().numpy().copy(), result.gradient_labels(b).cpu().numpy().copy(),
                       result.values(b).cpu().numpy().copy(), result.gradients(b).cpu().numpy().copy()))
    return result.keys.copy(), blocks


def assert_same(got, expected, what):
    keys_g, blocks_g = got
    keys_e, blocks_e = expected
    assert np.array_equal(keys_g, keys_e), what
    assert len(blocks_g) == len(blocks_e), what
    for b, (g, e) in enumerate(zip(blocks_g, blocks_e)):
        for name, x, y in zip(("sample labels", "gradient labels", "values", "gradients"), g, e):
            assert x.shape == y.shape, f"{what}: block {b} {name} shape"
            assert np.array_equal(x, y), f"{what}: block {b} {name}"


def clear(result):
    for b in range(result.n_blocks):
        result.values(b).fill_(float("nan"))
        result.gradients(b).fill_(float("nan"))
        result.sample_labels(b).fill_(-7)
        result.gradient_labels(b).fill_(-7)


def make_calculator(cls, hypers, precision):
    calc = cls(**hypers)
    calc.set_precision(precision)
    return calc


@pytest.mark.parametrize("kind,precision,n_lanes,n_batches", [
    ("power spectrum", 32, 3, 3),
    ("power spectrum", 32, 2, 5),
    ("spherical expansion", 64, 2, 3),
])
def test_replays_match_eager(kind, precision, n_lanes, n_batches):
    import torch

    cls, hypers = (ra.SoapPowerSpectrum, PS) if kind == "power spectrum" else (ra.SphericalExpansion, SE)
    batches = make_batches(n_batches, 6)
    say(kind, "precision", precision, "lanes", n_lanes, "batches", n_batches)
    # eager results of every batch, on a calculator outside the graph
    eager_calc = make_calculator(cls, hypers, precision)
    expected = []
    for b in batches:
        expected.append(snapshot(compute_device(eager_calc, b, ["positions"])))
    torch.cuda.synchronize()

    lanes = [make_calculator(cls, hypers, precision) for _ in range(n_lanes)]
    graph = BatchGraph(lanes, batches, ["positions"])
    try:
        resident = list(range(max(0, n_batches - n_lanes), n_batches))
        for replay in range(3):
            for k in resident:
                clear(graph.output(k))
            torch.cuda.synchronize()
            graph.launch()
            graph.check()
            for k in resident:
                assert_same(snapshot(graph.output(k)), expected[k], f"replay {replay}, batch {k}")
            # eager work between replays, on another calculator and other batches
            other = make_batches(1, 4, offset=40)[0]
            compute_device(eager_calc, other, ["positions"])
            torch.cuda.synchronize()
            say("replay", replay, "matches")
    finally:
        graph.close()


def test_bound_calculators_are_refused_until_closed():
    import torch

    batches = make_batches(2, 4)
    lanes = [make_calculator(ra.SoapPowerSpectrum, PS, 32) for _ in range(2)]
    graph = BatchGraph(lanes, batches, ["positions"])
    with pytest.raises(ra.FeatomicError, match="batch graph"):
        compute_device(lanes[0], batches[0], ["positions"])
    with pytest.raises(ra.FeatomicError, match="one batch graph"):
        BatchGraph([lanes[1]], batches, ["positions"])
    graph.launch()
    graph.check()
    graph.close()
    # released: the calculators work directly again
    compute_device(lanes[0], batches[0], ["positions"])
    torch.cuda.synchronize()


def test_graph_keeps_its_positions():
    """The graph copies the positions at creation: changing the caller's tensor
    afterwards does not change what a replay computes."""
    import torch

    batches = make_batches(2, 5)
    eager_calc = make_calculator(ra.SoapPowerSpectrum, PS, 32)
    expected = [snapshot(compute_device(eager_calc, b, ["positions"])) for b in batches]
    torch.cuda.synchronize()
    lanes = [make_calculator(ra.SoapPowerSpectrum, PS, 32) for _ in range(2)]
    graph = BatchGraph(lanes, batches, ["positions"])
    try:
        for b in batches:
            b.positions.mul_(1.5)
        torch.cuda.synchronize()
        graph.launch()
        graph.check()
        for k in range(2):
            assert_same(snapshot(graph.output(k)), expected[k], f"batch {k}")
    finally:
        graph.close()
