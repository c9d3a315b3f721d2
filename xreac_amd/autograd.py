"""The ReaxFF energy as a differentiable function of positions for PyTorch.

The reference obtains forces from HIPS Autograd's reverse pass over the
energy (``make_vjp``, src/xreac/calculator.py:141-148), with two custom
derivative rules (``scatter_sum``, neighbors.py:56-62; ``lammps_angle``,
energy.py:49-55).  Here the reverse pass is the library's hand-written adjoint
kernels, which run inside the same evaluation as the energy; this
``torch.autograd.Function`` exposes them to PyTorch: ``forward`` evaluates
energy and forces on the device, ``backward`` returns ``-forces * grad``.

The derivative follows the evaluation's force convention: fixed-charge forces
(charges held at their QEq values, LAMMPS angle rule) by default, or the
charge-response forces (the exact gradient of the QEq-equilibrated energy)
with ``full_derivative=True``.  Second derivatives are not available.
"""

from __future__ import annotations

import numpy as np

from .engine import Engine
from .geometry import Boundary, fit_device_range


def _function():
    import torch

    class EnergyFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, positions, potential):
            out = potential._evaluate(positions.detach())
            ctx.save_for_backward(out["forces"])
            potential.last = out
            return torch.tensor(float(out["components"].sum()), dtype=torch.float64, device=positions.device)

        @staticmethod
        def backward(ctx, grad):
            (forces,) = ctx.saved_tensors
            return -forces * grad.to(forces.device), None

    return EnergyFunction


_FUNCTION = None


class ReaxFFPotential:
    """``energy = potential(positions)`` for a fixed system (symbols, cell, charge).

    positions: (N, 3) float64 CUDA tensor in Angstrom; the result is a 0-d
    float64 tensor in kcal/mol whose ``backward`` fills ``positions.grad`` with
    minus the forces of the same evaluation.  ``last`` holds that evaluation's
    device outputs (forces, charges, components, counts).
    """

    def __init__(self, force_field, symbols, *, cell=None, pbc=None, total_charge=0.0, full_derivative=False,
                 qeq_tolerance=1e-10, precision="fp64", device=0, engine=None):
        import torch

        force_field.validate_model(tuple(symbols))
        self.torch = torch
        self.engine = engine if engine is not None else Engine(force_field, device)
        self.symbols = tuple(symbols)
        self.boundary = Boundary(cell, pbc)
        self.cell, self.pbc = cell, pbc
        self.total_charge = float(total_charge)
        self.full_derivative = bool(full_derivative)
        self.tol = float(qeq_tolerance)
        self.precision = precision
        self.cutoff = force_field.general[12]
        self.types = torch.as_tensor(self.engine.type_ids(self.symbols), device=self.engine.device)
        self.last = None

    def _evaluate(self, positions):
        torch = self.torch
        n = len(self.symbols)
        if positions.shape != (n, 3) or positions.dtype != torch.float64 or not positions.is_cuda:
            raise ValueError("positions must be an (N, 3) float64 tensor on the GPU")
        x = positions.contiguous()
        if self.boundary.periodic:
            # atoms far outside the cell are moved by whole cells (results unchanged)
            xh = x.cpu().numpy()
            moved, _ = fit_device_range(xh, self.cell, self.pbc, self.cutoff)
            if moved is not xh and not np.array_equal(moved, xh):
                x = torch.as_tensor(moved, device=positions.device)
        return self.engine.evaluate(
            x, self.types, cell=self.boundary.cell, pbc=self.boundary.pbc, total_charge=self.total_charge,
            tol=self.tol, lammps_angles=not self.full_derivative, full_derivative=self.full_derivative,
            properties=False, precision=self.precision,
        )

    def __call__(self, positions):
        global _FUNCTION
        if _FUNCTION is None:
            _FUNCTION = _function()
        return _FUNCTION.apply(positions, self)
