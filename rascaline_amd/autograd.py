"""Position gradients through torch's autograd, the role featomic-torch plays for
the reference (featomic-torch/src/autograd.cpp:135-230 for the backward, 237-300
for the contraction with dA/dX): the forward returns the calculator's values on the
GPU; the backward is the vector-Jacobian product of the position gradients
(`featomic_amd_positions_vjp`), computed without the gradient rows.

    positions = torch.tensor(..., dtype=torch.float64, device="cuda", requires_grad=True)
    values, layout = soap_values(calculator, systems, positions)   # flat values, block layout
    loss = (values * weights).sum()
    loss.backward()                                                 # positions.grad [atoms, 3]

`systems` is a list of (positions, types, cell) tuples giving the types and cells
(their positions are ignored); `positions` holds every atom's position, systems in
order. The values keep the calculator's precision (float32 or float64).
"""

import torch

from .device import DeviceSystems, compute_device, positions_vjp


class BlockLayout:
    """Where each block's values sit in the flat value tensor."""

    def __init__(self, result):
        self.keys = result.keys
        self.n_samples = list(result.n_samples)
        self.properties = list(result.properties)
        self.offsets = [result.value_offset(b) for b in range(result.n_blocks)]

    def block(self, values, b):
        """Block b of the flat values as [samples, properties] (a view)."""
        off, ns, npr = self.offsets[b], self.n_samples[b], self.properties[b]
        return values[off:off + ns * npr].view(ns, npr)


class _SoapValues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, calculator, systems):
        ds = DeviceSystems(systems, positions=positions.detach().contiguous())
        result = compute_device(calculator, ds)
        values = torch.empty(result.value_size, dtype=torch.float32 if result.precision == 32 else torch.float64,
                             device=positions.device)
        for b in range(result.n_blocks):
            off = result.value_offset(b)
            v = result.values(b)
            values[off:off + v.numel()] = v.reshape(-1)
        # saved for torch's in-place modification check; the backward reads the same storage
        ctx.save_for_backward(positions)
        ctx.calculator = calculator
        ctx.systems = ds
        ctx.layout = BlockLayout(result)
        return values

    @staticmethod
    def backward(ctx, grad_values):
        ctx.saved_tensors  # raises if the positions were modified in place since the forward
        grad = positions_vjp(ctx.calculator, ctx.systems, grad_values.contiguous())
        return grad.to(torch.float64), None, None


def soap_values(calculator, systems, positions):
    """Values of a calculator (SOAP spherical expansion or power spectrum, LODE
    spherical expansion) as one flat tensor plus its BlockLayout; differentiable with
    respect to `positions`."""
    values = _SoapValues.apply(positions, calculator, systems)
    # the layout of this computation (the same batch gives the same layout)
    layout = values.grad_fn.layout if values.grad_fn is not None else None
    if layout is None:
        ds = DeviceSystems(systems, positions=positions.detach().contiguous())
        layout = BlockLayout(compute_device(calculator, ds))
    return values, layout
