"""Differentiable degridding: image -> model visibilities as a torch.autograd.Function.

Forward (the reference's forward operator, StandardMfsModelPredictor /
residual refresh, lib.rs:3375-3404 and gridder.rs:1278-1312 / :3726-3747):
    image --apodize, centered FFT--> model grid --degrid--> visibilities.
Backward is the hand-written adjoint: the incoming visibility gradient is
gridded with unit weight (chip_vset_grid_values, the exact adjoint of the
degridder) and mapped back through the adjoint of the model transform
(chip_grid_to_model_adjoint).  No torch op is differentiated: both directions
run the HIP kernels.

For a real image x and complex visibilities V = A x, torch passes the
conjugate-Wirtinger gradient g of a real loss; the image gradient is Re(A^H g).
"""
from __future__ import annotations

import torch

from .imaging import StandardGridder, VisibilitySet


class DegridOperator:
    """Binds a gridder and a resident visibility set; ``op(image)`` is differentiable."""

    def __init__(self, gridder: StandardGridder, vset: VisibilitySet):
        if vset.n_chan != 1:
            raise ValueError("the differentiable operator takes one channel")
        self.gridder = gridder
        self.vset = vset

    def forward_linear(self, image: torch.Tensor) -> torch.Tensor:
        grid = self.gridder.model_to_grid(image)
        return self.vset.degrid(grid)

    def adjoint_linear(self, vis_grad: torch.Tensor) -> torch.Tensor:
        grid = self.gridder.new_grids(1)
        self.vset.grid_values(vis_grad, grid)
        return self.gridder.grid_to_model_adjoint(grid)[0]

    def __call__(self, image: torch.Tensor) -> torch.Tensor:
        return _Degrid.apply(image, self)


class _Degrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, op):
        ctx.op = op
        return op.forward_linear(image.detach())

    @staticmethod
    def backward(ctx, grad_vis):
        op = ctx.op
        g = op.adjoint_linear(grad_vis.contiguous())
        return g, None
