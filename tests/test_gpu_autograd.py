"""The energy as a torch.autograd.Function over the adjoint kernels (xreac_amd.autograd).

``energy.backward()`` must leave exactly minus the forces of ``Calculator.evaluate``
in ``positions.grad`` (the reference's make_vjp pass, calculator.py:141-148);
with charge-response forces the gradient must match central differences of the
energy.
"""

import numpy as np
import pytest

from tests import cases
from xreac_amd import Calculator, ForceField, workloads
from xreac_amd.autograd import ReaxFFPotential

pytestmark = pytest.mark.gpu


def _positions(x):
    import torch

    return torch.tensor(np.asarray(x, dtype=np.float64), device="cuda", requires_grad=True)


@pytest.mark.parametrize("full_derivative", [False, True])
def test_backward_is_minus_forces(full_derivative):
    ff = ForceField.bundled("ffield.reax.HO.2015")
    s, x, cell = workloads.water_box(side=4, seed=3)
    pot = ReaxFFPotential(ff, s, cell=cell, full_derivative=full_derivative, qeq_tolerance=1e-12)
    xt = _positions(x)
    energy = pot(xt)
    assert energy.requires_grad and energy.dtype == xt.dtype and energy.device == xt.device
    energy.backward()
    ref = Calculator(ff, qeq_tolerance=1e-12).evaluate(s, x, cell=cell, full_derivative=full_derivative)
    assert energy.item() == ref.energy
    np.testing.assert_array_equal(xt.grad.cpu().numpy(), -ref.forces)
    # chain rule through an outer scale factor
    xt.grad = None
    (2.5 * pot(xt)).backward()
    np.testing.assert_array_equal(xt.grad.cpu().numpy(), -2.5 * ref.forces)


def test_exact_gradient_matches_central_differences():
    ff = ForceField.bundled("ffield.reax.ZnOH.2010")
    s, x = cases.ZNO["cluster20"]
    pot = ReaxFFPotential(ff, s, full_derivative=True, qeq_tolerance=1e-12)
    xt = _positions(x)
    pot(xt).backward()
    grad = xt.grad.cpu().numpy()
    rng = np.random.default_rng(5)
    h = 1e-5
    for _ in range(6):
        d = rng.normal(size=x.shape)
        d /= np.linalg.norm(d)
        ep = pot(_positions(x + h * d)).item()
        em = pot(_positions(x - h * d)).item()
        assert (ep - em) / (2 * h) == pytest.approx(np.sum(grad * d), abs=2e-5, rel=1e-6)


def test_optimizer_steps_lower_the_energy():
    import torch

    ff = ForceField.bundled("ffield.reax.HO.2015")
    s, x = workloads.water_cluster(molecules=12, radius=5.0, seed=2)
    pot = ReaxFFPotential(ff, s)
    xt = _positions(x)
    opt = torch.optim.SGD([xt], lr=1e-4)
    energies = []
    for _ in range(5):
        opt.zero_grad()
        e = pot(xt)
        e.backward()
        energies.append(e.item())
        with torch.no_grad():
            xt.grad.clamp_(-500.0, 500.0)
        opt.step()
    assert energies[-1] < energies[0]
