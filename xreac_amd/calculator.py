"""Public calculator: the reference's ``Calculator`` API on the GPU path.

Signatures, defaults, result types and error texts follow
``src/xreac/calculator.py:15-246`` (hsulab/xreac 0.10.0).  The work of
``evaluate`` runs in ``libxreac_hip.so``; this module only validates inputs,
moves arrays and assembles the ``Evaluation``.
"""

from __future__ import annotations

from dataclasses import dataclass
from numbers import Real

import numpy as np

from .engine import COMPONENTS, Engine
from .geometry import Boundary, canonical_neighbors, fit_device_range, validate_expansion_limit


@dataclass(frozen=True)
class Evaluation:
    energy: float
    forces: np.ndarray
    charges: np.ndarray
    components: dict
    full_derivative: bool
    bond_orders: np.ndarray
    total_bond_orders: np.ndarray
    lone_pairs: np.ndarray
    bond_counts: np.ndarray
    dipole: np.ndarray
    cell_repetitions: tuple = (1, 1, 1)
    neighbor_backend: str = "replicated"

    @property
    def force_convention(self):
        """Name of the derivative used for ``forces``."""
        return "charge_response" if self.full_derivative else "fixed_charge"


@dataclass(frozen=True)
class Relaxation:
    positions: np.ndarray
    evaluation: Evaluation
    converged: bool
    iterations: int
    message: str


def validate_total_charge(total_charge):
    """Finite real scalar in e, fractions allowed (calculator.py:45-52)."""
    if isinstance(total_charge, (bool, np.bool_)) or not isinstance(total_charge, Real) or not np.isfinite(total_charge):
        raise ValueError("total_charge must be a finite real scalar")


def validate_input(symbols, positions, total_charge=0, cell=None, pbc=None):
    """Host-side checks of calculator.py:55-66.  Coincident atoms are detected on the device."""
    symbols = tuple(symbols)
    if not symbols or any(not isinstance(s, str) or not s for s in symbols):
        raise ValueError("A nonempty sequence of parameter-file atom labels is required")
    validate_total_charge(total_charge)
    x = np.array(positions, dtype=np.float64, copy=True)
    if x.shape != (len(symbols), 3) or not np.isfinite(x).all():
        raise ValueError("positions must be a finite (N, 3) array in Angstrom")
    Boundary(cell, pbc)
    return symbols, x


class Calculator:
    def __init__(self, force_field, *, max_expanded_atoms=512, device=0, qeq_tolerance=1e-10):
        force_field.validate_model()
        validate_expansion_limit(max_expanded_atoms)
        self.force_field = force_field
        self.max_expanded_atoms = max_expanded_atoms
        self.qeq_tolerance = float(qeq_tolerance)
        self._device = device
        self._engine = None

    @property
    def engine(self):
        if self._engine is None:
            self._engine = Engine(self.force_field, self._device)
        return self._engine

    def evaluate(self, symbols, positions, *, total_charge=0, cell=None, pbc=None, full_derivative=False,
                 neighbors=None):
        """Energy, QEq charges and forces on the GPU (calculator.py:76-165).

        Fixed-charge forces with the LAMMPS angle regularisation by default;
        full_derivative=True gives charge-response forces (one adjoint QEq
        solve) with the plain angle derivative, as the reference does.
        """
        if not isinstance(full_derivative, bool):
            raise ValueError("full_derivative must be a boolean")
        symbols, x = validate_input(symbols, positions, total_charge, cell, pbc)
        self.force_field.validate_model(symbols)
        boundary = Boundary(cell, pbc)
        csr = None
        if neighbors is not None:
            # the caller's full directed list (calculator.py:113-115): validated and put in canonical order
            csr = canonical_neighbors(neighbors, len(x), cell, pbc)
            repetitions, backend = (1, 1, 1), "provided"
        else:
            repetitions = boundary.repetitions(self.force_field.general[12], len(x), self.max_expanded_atoms)
            backend = "replicated"
        # whole-cell moves of far atoms and of huge supplied shifts (results unchanged)
        x_dev, csr = fit_device_range(x, cell, pbc, self.force_field.general[12], csr)
        eng = self.engine
        torch = eng.torch
        types = torch.as_tensor(eng.type_ids(symbols), device=eng.device)
        xd = torch.as_tensor(x_dev, device=eng.device)
        out = eng.evaluate(
            xd,
            types,
            cell=boundary.cell,
            pbc=boundary.pbc,
            total_charge=total_charge,
            tol=self.qeq_tolerance,
            lammps_angles=not full_derivative,
            full_derivative=full_derivative,
            neighbors=csr,
        )
        comps = out["components"]
        forces = out["forces"].cpu().numpy()
        charges = out["charges"].cpu().numpy()
        if not all(np.isfinite(v).all() for v in (comps, charges, forces)):
            raise ValueError("Non-finite energy, charges, or forces; check the geometry")
        if abs(charges.sum() - total_charge) > 1e-8:
            raise ValueError("QEq charge constraint failed")
        n = len(x)
        bi, bj, _, bo = eng.bonds(out["num_bonds"])
        dense = np.zeros((n, n))
        np.add.at(dense, (bi, bj), bo)
        masses = np.array([self.force_field.atoms[s]["mass"] for s in symbols])
        center = np.sum(x * masses[:, None], axis=0) / np.sum(masses)
        return Evaluation(
            float(comps.sum()),
            forces,
            charges,
            dict(zip(COMPONENTS, map(float, comps))),
            full_derivative,
            bond_orders=dense,
            total_bond_orders=out["total_bond_orders"].cpu().numpy(),
            lone_pairs=out["lone_pairs"].cpu().numpy(),
            bond_counts=out["bond_counts"].cpu().numpy().astype(int),
            dipole=np.sum((x - center) * charges[:, None], axis=0),
            cell_repetitions=repetitions,
            neighbor_backend=backend,
        )

    def relax(self, symbols, positions, *, force_tolerance=1e-4, max_iterations=500, total_charge=0, cell=None,
              pbc=None, backend="ase"):
        """Fixed-charge FIRE relaxation (calculator.py:167-246); only the native backend is available."""
        symbols, x = validate_input(symbols, positions, total_charge, cell, pbc)
        if not np.isfinite(force_tolerance) or force_tolerance <= 0:
            raise ValueError("force_tolerance must be positive and finite")
        if isinstance(max_iterations, bool) or not isinstance(max_iterations, int) or max_iterations < 1:
            raise ValueError("max_iterations must be a positive integer")
        if backend not in ("ase", "native"):
            raise ValueError("backend must be 'ase' or 'native'")
        if backend == "ase":
            raise ImportError("the ASE backend needs the ase package, which is not installed; use backend='native'")
        # FIRE on the device (xreac_fire_step restates calculator.py:225-245): positions, velocities and
        # the QEq state stay in HBM; only the per-iteration status word comes back.
        eng = self.engine
        torch = eng.torch
        boundary = Boundary(cell, pbc)
        n = len(x)
        xd = torch.as_tensor(x, device=eng.device).contiguous()
        types = torch.as_tensor(eng.type_ids(symbols), device=eng.device)
        velocity = torch.zeros_like(xd)
        state = torch.tensor([[0.02, 0.1, 0.0, 0.0]], dtype=torch.float64, device=eng.device)
        status = torch.zeros(1, dtype=torch.int32, device=eng.device)
        offsets = np.array([0, n], dtype=np.int32)
        off_dev = torch.as_tensor(offsets, device=eng.device)
        periodic = boundary.cell is not None and boundary.periodic
        cells = (np.asarray(boundary.cell, dtype=np.float64) if periodic else np.eye(3)).reshape(1, 3, 3)
        pbcs = (np.asarray(boundary.pbc, dtype=np.int32) if periodic else np.zeros(3, dtype=np.int32)).reshape(1, 3)
        out = eng.batch_outputs(n, qeq_state=True)
        code = 0
        for it in range(max_iterations + 1):
            res = eng.evaluate_batch(xd, types, offsets, cells, pbcs, [float(total_charge)], tol=self.qeq_tolerance,
                                     warm_start=it > 0, out=out)
            if res["errors"][0] is not None:
                raise ValueError(res["errors"][0])
            eng.fire_step(off_dev, xd, velocity, out["forces"], state, status, force_tolerance, max_iterations)
            code = int(status.item())
            if code:
                break
        x = xd.cpu().numpy()
        iterations = int(state[0, 3].item())
        result = self.evaluate(symbols, x, total_charge=total_charge, cell=cell, pbc=pbc)
        if code == 1:
            return Relaxation(x, result, True, iterations, "Fixed-charge force tolerance reached")
        return Relaxation(x, result, False, iterations, "Maximum relaxation iterations reached")
