"""torch.autograd.Function over the C-ABI: the snapshot temperature field as a
differentiable function of the beam parameters and the path vertices.

    field = SnapshotField(thesis)                  # a single-beam Thesis context
    T = field(beam5, xy)                           # (pnum,) float64 CUDA tensor
    loss = ((T - T_target) ** 2).mean(); loss.backward()
    beam5.grad, xy.grad                            # dL/d(Power, Efficiency, Width_X, Width_Y, Depth_Z), dL/d(x, y) mm

The forward pass is the FP32-compensated (or FP64) node sum on the GPU; the
backward pass is the hand-written adjoint kernel plus the host chain through the
quadrature schedule (branch decisions frozen). There is no autograd tape through
the kernels themselves.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import FP32C, Thesis


class _SnapshotFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, beam5: torch.Tensor, xy: torch.Tensor, field: "SnapshotField") -> torch.Tensor:
        rows = field.rows_template.copy()
        rows[:, 1:3] = xy.detach().to("cpu", torch.float64).numpy()
        beam = beam5.detach().to("cpu", torch.float64).numpy().copy()
        th = field.thesis
        th.set_design(beam, rows)
        T = torch.empty(th.info.pnum, dtype=torch.float64, device="cuda")
        # the library's entry points take 0 for "the context's own stream",
        # which torch's default stream also reports: order the two explicitly
        # (T is read by torch as soon as this returns)
        torch.cuda.current_stream().synchronize()
        th.forward_dev(T.data_ptr(), field.precision, field.t)
        torch.cuda.synchronize()  # every stream of the device, the context's included
        ctx.field = field
        ctx.design = (beam, rows)
        ctx.dev = beam5.device, xy.device
        return T

    @staticmethod
    def backward(ctx, grad_T: torch.Tensor):
        field = ctx.field
        th = field.thesis
        th.set_design(*ctx.design)
        gT = grad_T.detach().to(torch.float64).contiguous()
        torch.cuda.current_stream().synchronize()
        g = th.adjoint_dev(gT.data_ptr(), field.t)
        m = ctx.design[1].shape[0]
        g_beam = torch.from_numpy(g[:5].copy()).to(ctx.dev[0])
        g_xy = torch.from_numpy(g[5:].reshape(m, 2).copy()).to(ctx.dev[1])
        return g_beam, g_xy, None


class SnapshotField(torch.nn.Module):
    """T at every grid point of a single-beam deck, differentiable in the beam
    parameters and the x/y of every Path-file row (file units, mm)."""

    def __init__(self, thesis: Thesis, precision: int = FP32C, t: float = -1.0,
                 rows_template: Optional[np.ndarray] = None):
        super().__init__()
        self.thesis = thesis
        self.precision = precision
        self.t = t
        if rows_template is None:
            rows_template = thesis.design()[1]
        self.rows_template = np.array(rows_template, dtype=np.float64)

    def forward(self, beam5: torch.Tensor, xy: torch.Tensor) -> torch.Tensor:
        return _SnapshotFn.apply(beam5, xy, self)
