"""torch.autograd bridge over the adjoint pair of the C ABI (vfvm_adjoint_forward /
vfvm_adjoint_backward): the final state of the linear steady or implicit-Euler map as a
differentiable function of the per-vertex values of one scalar field.

The library runs its own HIP runtime; torch's ROCm runtime initialises only if torch is
imported before the first vfvm.Context is created, so import torch first when the
tensors live on the GPU."""
from __future__ import annotations

import numpy as np
import torch

from .core import WRT_SCLDOM, Context


class FieldToState(torch.autograd.Function):
    """T^num_ts = F(values of field `field_id`).

    forward: uploads the field values, resets the unknown to `initial`, runs the forward
    pass (checkpoints every `every` steps) and returns T^num_ts.
    backward: runs the backward pass with seed = grad_output and returns dJ/dvalues.
    Tensors are FP64; their device may be the CPU or the GPU (values cross through host
    memory, as the C ABI's field set/get do)."""

    @staticmethod
    def forward(fctx, values: torch.Tensor, vctx: Context, field_id: int, ops, initial: np.ndarray, dt: float,
                num_ts: int, every: int, lin: dict | None, lin_adj: dict | None):
        unk = ops[0][2]
        vctx.scldom_set(field_id, values.detach().to("cpu", torch.float64).numpy())
        vctx.scldom_set(unk, initial)
        vctx.adjoint_forward(ops, dt=dt, num_ts=num_ts, checkpoint_every=every, lin=lin)
        out = torch.from_numpy(vctx.scldom_get(unk)).to(values.device)
        fctx.vctx, fctx.field_id, fctx.lin_adj = vctx, field_id, lin_adj
        return out

    @staticmethod
    def backward(fctx, grad_out: torch.Tensor):
        seed = grad_out.detach().to("cpu", torch.float64).numpy()
        (g,), _ = fctx.vctx.adjoint_backward(seed, [(WRT_SCLDOM, fctx.field_id)], lin=fctx.lin_adj)
        return (torch.from_numpy(g).to(grad_out.device),) + (None,) * 9


def field_to_state(values, vctx, field_id, ops, initial, dt=0.0, num_ts=0, every=10, lin=None, lin_adj=None):
    """Differentiable T^num_ts(values); see FieldToState."""
    return FieldToState.apply(values, vctx, field_id, ops, np.ascontiguousarray(initial, np.float64), float(dt),
                              int(num_ts), int(every), lin, lin_adj)
