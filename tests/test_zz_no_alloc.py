"""The allocation-free training pass (featomic_amd_mlp_loss_gradient_noalloc, device
arrays, scratch in a static device array) gives the same loss and gradient as
featomic_amd_mlp_loss_gradient (host arrays, per-call device buffers), bit for bit:
both run the same kernels over the same tiles. Two calls in a row, the second with
a smaller input, so that the second reads a scratch holding the first call's data;
and a call beyond the scratch's capacity is refused."""

import ctypes
import sys

import numpy as np
import pytest

from rascaline_amd import _c_lib as C

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def n_params(n_out):
    return 64 + 64 + 64 * 64 + 64 + n_out * 64 + n_out


def host_pass(n_out, params, inputs, targets):
    loss = ctypes.c_double()
    gradient = np.zeros(n_params(n_out), dtype=np.float32)
    C.check(C.lib().featomic_amd_mlp_loss_gradient(n_out, len(inputs), params.ctypes.data, inputs.ctypes.data,
                                                   targets.ctypes.data, ctypes.byref(loss), gradient.ctypes.data))
    return loss.value, gradient


def device_pass(n_out, params, inputs, targets):
    import torch

    p = torch.from_numpy(params).cuda()
    x = torch.from_numpy(inputs).cuda()
    t = torch.from_numpy(targets).cuda()
    loss = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    gradient = torch.full((n_params(n_out),), float("nan"), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    C.check(C.lib().featomic_amd_mlp_loss_gradient_noalloc(n_out, len(inputs), p.data_ptr(), x.data_ptr(), t.data_ptr(),
                                                           loss.data_ptr(), gradient.data_ptr(),
                                                           ctypes.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return float(loss.cpu()[0]), gradient.cpu().numpy()


def case(rng, n_out, n_samples):
    params = (rng.normal(size=n_params(n_out)) * 0.2).astype(np.float32)
    inputs = rng.uniform(-1, 1, n_samples).astype(np.float32)
    targets = rng.normal(size=(n_samples, n_out)).astype(np.float32)
    return params, inputs, targets


@pytest.mark.parametrize("n_out", [216, 6])
def test_no_alloc_pass_matches_the_allocating_one(n_out):
    import torch

    torch.cuda.init()
    rng = np.random.default_rng(n_out)
    for n_samples in (3000, 700):
        params, inputs, targets = case(rng, n_out, n_samples)
        say("no-alloc loss gradient, n_out", n_out, "samples", n_samples)
        expected = host_pass(n_out, params, inputs, targets)
        got = device_pass(n_out, params, inputs, targets)
        assert got[0] == expected[0], (got[0], expected[0])
        assert np.array_equal(got[1], expected[1]), float(np.abs(got[1] - expected[1]).max())


def test_no_alloc_pass_refuses_inputs_beyond_its_scratch():
    import torch

    torch.cuda.init()
    n_out = 216
    rng = np.random.default_rng(1)
    params, inputs, targets = case(rng, n_out, 40000)
    with pytest.raises(C.FeatomicError, match="at most"):
        device_pass(n_out, params, inputs, targets)
