"""Pin the oracle's radial-surrogate network (forward, loss and loss gradient)
independently of its own formulas: central finite differences of the loss, and
torch's float64 automatic differentiation of the same network written with
torch.nn layers. A dropped backward term, a transposed weight or a wrong SiLU
derivative fails both."""

import numpy as np
import pytest
import torch

import featomic_oracle as orc

H = orc.MLP_HIDDEN


def random_params(n_out, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return scale * rng.normal(size=orc.mlp_n_params(n_out))


def torch_network(params, n_out):
    """The same network as torch.nn layers, float64, parameters copied from the
    flat layout W1[64], b1[64], W2[64][64], b2[64], W3[n_out][64], b3[n_out]."""
    net = torch.nn.Sequential(torch.nn.Linear(1, H), torch.nn.SiLU(), torch.nn.Linear(H, H), torch.nn.SiLU(),
                              torch.nn.Linear(H, n_out)).double()
    p = torch.tensor(params, dtype=torch.float64)
    sizes = [H, H, H * H, H, n_out * H, n_out]
    chunks = torch.split(p, sizes)
    with torch.no_grad():
        net[0].weight.copy_(chunks[0].reshape(H, 1))
        net[0].bias.copy_(chunks[1])
        net[2].weight.copy_(chunks[2].reshape(H, H))
        net[2].bias.copy_(chunks[3])
        net[4].weight.copy_(chunks[4].reshape(n_out, H))
        net[4].bias.copy_(chunks[5])
    return net


def torch_gradient(net):
    return torch.cat([net[0].weight.grad.reshape(-1), net[0].bias.grad, net[2].weight.grad.reshape(-1),
                      net[2].bias.grad, net[4].weight.grad.reshape(-1), net[4].bias.grad]).numpy()


@pytest.mark.parametrize("n_out", [1, 6, 216])
def test_mlp_forward_matches_torch(n_out):
    params = random_params(n_out, 1)
    x = np.linspace(-1.0, 1.0, 17)
    y = orc.mlp_forward(params, x, n_out)
    with torch.no_grad():
        expected = torch_network(params, n_out)(torch.tensor(x)[:, None]).numpy()
    np.testing.assert_allclose(y, expected, rtol=1e-13, atol=1e-13 * np.abs(expected).max())


@pytest.mark.parametrize("n_out", [3, 216])
def test_mlp_loss_gradient_matches_torch_autograd(n_out):
    rng = np.random.default_rng(2)
    params = random_params(n_out, 3)
    x = rng.uniform(-1, 1, 11)
    t = rng.normal(size=(11, n_out))
    loss, grad = orc.mlp_loss_gradient(params, x, t, n_out)
    net = torch_network(params, n_out)
    out = net(torch.tensor(x)[:, None])
    ref = ((out - torch.tensor(t)) ** 2).mean()
    ref.backward()
    assert abs(loss - ref.item()) <= 1e-13 * abs(ref.item())
    expected = torch_gradient(net)
    np.testing.assert_allclose(grad, expected, rtol=1e-11, atol=1e-13 * np.abs(expected).max())


def test_mlp_loss_gradient_finite_differences():
    n_out = 5
    rng = np.random.default_rng(4)
    params = random_params(n_out, 5)
    x = rng.uniform(-1, 1, 7)
    t = rng.normal(size=(7, n_out))
    _, grad = orc.mlp_loss_gradient(params, x, t, n_out)
    # one parameter from every tensor, plus random ones
    sizes = np.cumsum([0, H, H, H * H, H, n_out * H, n_out])
    picks = list(sizes[:-1] + 3) + list(rng.integers(0, params.size, 20))
    h = 1e-6
    for k in picks:
        p1, p2 = params.copy(), params.copy()
        p1[k] += h
        p2[k] -= h
        fd = (orc.mlp_loss_gradient(p1, x, t, n_out)[0] - orc.mlp_loss_gradient(p2, x, t, n_out)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-7 * max(1.0, abs(grad[k])), (k, fd, grad[k])


def test_mlp_zero_residual_has_zero_gradient():
    n_out = 4
    params = random_params(n_out, 6)
    x = np.linspace(-1, 1, 9)
    t = orc.mlp_forward(params, x, n_out)
    loss, grad = orc.mlp_loss_gradient(params, x, t, n_out)
    assert loss == 0.0
    assert np.all(grad == 0.0)
