"""Position gradients through torch's autograd, the role featomic-torch plays for
the reference (featomic-torch/src/autograd.cpp:135-230 for the backward, 237-420
for the contraction with dA/dX and its own backward): the forward returns the
calculator's values on the GPU; the backward is the vector-Jacobian product of the
position gradients (`featomic_amd_positions_vjp`), computed without the gradient
rows; and that backward is itself differentiable with respect to dA/dX, as
featomic-torch's PositionsGrad::backward is (autograd.cpp:320-420), so that a loss
on forces (create_graph=True) reaches the weights applied to the descriptor.

    positions = torch.tensor(..., dtype=torch.float64, device="cuda", requires_grad=True)
    values, layout = soap_values(calculator, systems, positions)   # flat values, block layout
    energy = (values * weights).sum()
    forces = -torch.autograd.grad(energy, positions, create_graph=True)[0]
    loss = ((forces - reference) ** 2).sum()
    loss.backward()                                                 # weights.grad

`systems` is a list of (positions, types, cell) tuples giving the types and cells
(their positions are ignored); `positions` holds every atom's position, systems in
order. The values keep the calculator's precision (float32 or float64).

With `cell` ([systems][3][3] float64) the values are also differentiable with
respect to the cells, at fixed Cartesian positions, as featomic-torch's CellGrad
makes them (featomic-torch/src/autograd.cpp:425-560): the forward keeps the
calculator's cell gradients dX/dH (SOAP calculators), the backward contracts them
with dA/dX per system, and that backward is differentiable with respect to dA/dX.

As in featomic-torch, second derivatives with respect to the positions or the cell
are not computed: the double backward contributes nothing to them and warns once.
"""

import warnings

import numpy as np
import torch

from .device import DeviceSystems, compute_device, positions_vjp


class BlockLayout:
    """Where each block's values sit in the flat value tensor."""

    def __init__(self, result):
        self.keys = result.keys
        self.n_samples = list(result.n_samples)
        self.properties = list(result.properties)
        self.offsets = [result.value_offset(b) for b in range(result.n_blocks)]
        self.value_size = result.value_size

    def block(self, values, b):
        """Block b of the flat values as [samples, properties] (a view)."""
        off, ns, npr = self.offsets[b], self.n_samples[b], self.properties[b]
        return values[off:off + ns * npr].view(ns, npr)


class _Holder:
    layout = None


def _lode_row_labels(systems, key):
    """(sample, system, atom) of the position-gradient rows of one LODE block, in the
    reference's order (labels/samples/long_range.rs:55-77): for every system holding the
    neighbour type, every center of the center type, the atoms of the neighbour type and
    the center itself."""
    _, _, ct, nt = (int(k) for k in key)
    rows = []
    sample = 0
    for s, (_, types, _) in enumerate(systems):
        types = np.asarray(types)
        if not np.any(types == nt):
            continue
        for i in np.nonzero(types == ct)[0]:
            atoms = np.nonzero((types == nt) | (np.arange(len(types)) == i))[0]
            rows.append(np.stack([np.full(len(atoms), sample), np.full(len(atoms), s), atoms], axis=1))
            sample += 1
    return np.concatenate(rows) if rows else np.zeros((0, 3), dtype=np.int64)


def positions_jvp(calculator, systems, device_systems, tangent):
    """dX/dr · v: for every sample, Σ over its position-gradient rows of (dX/dr)[row] ·
    tangent[atom of row], in the layout of the flat values (the product featomic-torch's
    PositionsGrad::backward forms, featomic-torch/src/autograd.cpp:380-410, from the
    calculator's own gradient rows, here on the device)."""
    result = compute_device(calculator, device_systems, ["positions"])
    dtype = torch.float32 if result.precision == 32 else torch.float64
    out = torch.zeros(result.value_size, dtype=dtype, device=tangent.device)
    starts = torch.from_numpy(device_systems.offsets[:-1].copy()).to(tangent.device)
    tangent = tangent.to(dtype)
    for b in range(result.n_blocks):
        g = result.gradients(b)
        if g.shape[0] == 0:
            continue
        labels = result.gradient_labels(b)
        if labels is None:
            labels = torch.from_numpy(_lode_row_labels(systems, result.keys[b])).to(tangent.device)
        labels = labels.long()
        atoms = starts[labels[:, 1]] + labels[:, 2]
        t = tangent[atoms]
        contribution = g[:, 0] * t[:, 0:1] + g[:, 1] * t[:, 1:2] + g[:, 2] * t[:, 2:3]
        # rows are sorted by sample: per-sample sums as differences of an FP64 running
        # sum at the segment ends (a fixed order, so repeated calls agree bit for bit)
        counts = torch.bincount(labels[:, 0], minlength=result.n_samples[b])
        ends = counts.cumsum(0)
        running = torch.cat([torch.zeros((1, contribution.shape[1]), dtype=torch.float64, device=tangent.device),
                             contribution.double().cumsum(0)])
        block = (running[ends] - running[ends - counts]).to(dtype)
        off = result.value_offset(b)
        out[off:off + block.numel()] = block.reshape(-1)
    return out


_WARNED = [False]
_WARNED_CELL = [False]


def _segment_sums(per_row, counts):
    """Sums of consecutive runs of rows ([rows, ...] -> [len(counts), ...]) as differences
    of an FP64 running sum at the run ends: a fixed order, so repeated calls agree bit
    for bit."""
    import torch

    ends = counts.cumsum(0)
    flat = per_row.reshape(per_row.shape[0], -1).double()
    running = torch.cat([torch.zeros((1, flat.shape[1]), dtype=torch.float64, device=flat.device), flat.cumsum(0)])
    return (running[ends] - running[ends - counts]).reshape((len(counts),) + tuple(per_row.shape[1:]))


class _CellBlocks:
    """dX/dH of every block ([samples, 3, 3, properties]) with the system of each sample
    (samples are sorted by system)."""

    def __init__(self, result, n_systems, device):
        import torch

        self.n_systems = n_systems
        self.blocks = []
        for b in range(result.n_blocks):
            ns, npr = result.n_samples[b], result.properties[b]
            if ns == 0:
                continue
            grad = result.cell_gradients(b).double().clone()
            systems = result.sample_labels(b)[:, 0].long().clone()
            counts = torch.bincount(systems, minlength=n_systems)
            self.blocks.append((result.value_offset(b), ns, npr, grad, counts))


class _CellGrad(torch.autograd.Function):
    """dA/dH from dA/dX (featomic-torch's CellGrad, autograd.cpp:428-489): per system,
    Σ over its samples and properties of dA/dX[sample] · dX/dH[sample]. Its backward gives
    dB/d(dA/dX)[sample] = dB/d(dA/dH)[system of sample] : dX/dH[sample] (:493-560)."""

    @staticmethod
    def forward(ctx, cell, grad_values, cell_blocks):
        ctx.save_for_backward(cell, grad_values)
        ctx.cell_blocks = cell_blocks
        out = torch.zeros((cell_blocks.n_systems, 3, 3), dtype=torch.float64, device=grad_values.device)
        for off, ns, npr, grad, counts in cell_blocks.blocks:
            w = grad_values[off:off + ns * npr].reshape(ns, 1, 1, npr).double()
            per_sample = (grad * w).sum(dim=3)  # [samples, 3, 3]
            out += _segment_sums(per_sample, counts)
        return out.to(cell.dtype)

    @staticmethod
    def backward(ctx, grad_grad):
        cell, grad_values = ctx.saved_tensors
        if ctx.needs_input_grad[0] and not _WARNED_CELL[0]:
            _WARNED_CELL[0] = True
            warnings.warn("second derivatives with respect to cell matrix are not implemented and are not "
                          "accumulated during backward() calls", stacklevel=2)
        d_grad_values = None
        if ctx.needs_input_grad[1]:
            d_grad_values = torch.zeros_like(grad_values)
            gg = grad_grad.double()
            for off, ns, npr, grad, counts in ctx.cell_blocks.blocks:
                systems = torch.repeat_interleave(torch.arange(len(counts), device=counts.device), counts)
                rows = (grad * gg[systems].reshape(ns, 3, 3, 1)).sum(dim=(1, 2))  # [samples, properties]
                d_grad_values[off:off + ns * npr] = rows.reshape(-1).to(grad_values.dtype)
        return None, d_grad_values, None


class _PositionsGrad(torch.autograd.Function):
    """dA/dr from dA/dX (featomic-torch's PositionsGrad); its backward gives dB/d(dA/dX)."""

    @staticmethod
    def forward(ctx, positions, grad_values, calculator, systems, device_systems):
        ctx.save_for_backward(positions, grad_values)
        ctx.calculator = calculator
        ctx.systems = systems
        ctx.device_systems = device_systems
        grad = positions_vjp(calculator, device_systems, grad_values.contiguous())
        return grad.to(torch.float64)

    @staticmethod
    def backward(ctx, grad_grad):
        positions, grad_values = ctx.saved_tensors
        if ctx.needs_input_grad[0] and not _WARNED[0]:
            _WARNED[0] = True
            warnings.warn("second derivatives with respect to positions are not implemented and are not "
                          "accumulated during backward() calls", stacklevel=2)
        d_grad_values = None
        if ctx.needs_input_grad[1]:
            d_grad_values = positions_jvp(ctx.calculator, ctx.systems, ctx.device_systems, grad_grad.contiguous())
            d_grad_values = d_grad_values.to(grad_values.dtype)
        return None, d_grad_values, None, None, None


class _SoapValues(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, cell, calculator, systems, holder):
        cells = None if cell is None else cell.detach().double().cpu().numpy().reshape(len(systems), 3, 3)
        ds = DeviceSystems(systems, positions=positions.detach().contiguous(), cells=cells)
        want_cell = cell is not None and cell.requires_grad
        result = compute_device(calculator, ds, ["cell"] if want_cell else [])
        values = torch.empty(result.value_size, dtype=torch.float32 if result.precision == 32 else torch.float64,
                             device=positions.device)
        for b in range(result.n_blocks):
            off = result.value_offset(b)
            v = result.values(b)
            values[off:off + v.numel()] = v.reshape(-1)
        holder.layout = BlockLayout(result)
        ctx.cell_blocks = _CellBlocks(result, len(systems), positions.device) if want_cell else None
        # saved for torch's in-place modification check; the backward reads the same storage
        ctx.save_for_backward(positions, cell)
        ctx.calculator = calculator
        ctx.systems = systems
        ctx.device_systems = ds
        return values

    @staticmethod
    def backward(ctx, grad_values):
        positions, cell = ctx.saved_tensors  # raises if they were modified in place since the forward
        grad_positions = grad_cell = None
        if ctx.needs_input_grad[0]:
            grad_positions = _PositionsGrad.apply(positions, grad_values, ctx.calculator, ctx.systems,
                                                  ctx.device_systems)
        if ctx.needs_input_grad[1]:
            grad_cell = _CellGrad.apply(cell, grad_values, ctx.cell_blocks).reshape(cell.shape)
        return grad_positions, grad_cell, None, None, None


def soap_values(calculator, systems, positions, cell=None):
    """Values of a calculator (SOAP spherical expansion or power spectrum, LODE
    spherical expansion) as one flat tensor plus its BlockLayout; differentiable with
    respect to `positions` and (SOAP, when given) `cell` ([systems][3][3] float64),
    twice with respect to what multiplies the values."""
    holder = _Holder()
    values = _SoapValues.apply(positions, cell, calculator, systems, holder)
    return values, holder.layout
