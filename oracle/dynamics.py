"""CPU restatements of the integrators -- test infrastructure only.

They check the device integrators (velocity Verlet and FIRE) with forces from
any callable, normally the FP64 oracle ``reaxff.evaluate``.

* ``fire`` follows the reference's native relaxation loop
  (src/xreac/calculator.py:213-246): unit fictitious masses, dt 0.02,
  dt_max 0.2, alpha 0.1, growth after more than five positive-power steps,
  0.1 A displacement cap.
* ``velocity_verlet`` is the textbook scheme (Swope, Andersen, Berens and
  Wilson, J. Chem. Phys. 76, 637 (1982)); the reference has no dynamics.
  Units: A, fs, amu, kcal/mol.

Pins (tests/test_oracle_pins.py): a harmonic oscillator (energy conservation
to O(dt^2) and the closed-form trajectory) and hand-derived first FIRE
iterations on a quadratic well.
"""

from __future__ import annotations

import numpy as np

ACCEL = 4.184e-4  # (kcal/mol/A/amu) -> A/fs^2


def velocity_verlet(force_fn, x, v, masses, dt, steps):
    """Integrate ``steps`` steps; returns (x, v, potential[steps+1], kinetic[steps+1]).

    force_fn(x) -> (energy, forces).
    """
    x = np.array(x, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    m = np.asarray(masses, dtype=np.float64)[:, None]
    e, f = force_fn(x)
    pot, kin = [e], [0.5 * np.sum(m * v * v) / ACCEL]
    for _ in range(steps):
        v = v + (0.5 * dt * ACCEL / m) * f
        x = x + dt * v
        e, f = force_fn(x)
        v = v + (0.5 * dt * ACCEL / m) * f
        pot.append(e)
        kin.append(0.5 * np.sum(m * v * v) / ACCEL)
    return x, v, np.array(pot), np.array(kin)


def fire(force_fn, x, force_tolerance, max_iterations):
    """Native FIRE (calculator.py:213-246); returns (x, converged, iterations, trajectory).

    force_fn(x) -> forces.  trajectory[k] holds the positions after k moves.
    """
    x = np.array(x, dtype=np.float64)
    velocity = np.zeros_like(x)
    dt, dt_max, alpha = 0.02, 0.2, 0.1
    positive = 0
    trajectory = [x.copy()]
    force = force_fn(x)
    for it in range(max_iterations + 1):
        if np.max(np.abs(force)) <= force_tolerance:
            return x, True, it, trajectory
        if it == max_iterations:
            break
        power = float(np.sum(velocity * force))
        if power > 0:
            velocity = (1 - alpha) * velocity + alpha * np.linalg.norm(velocity) / np.linalg.norm(force) * force
            positive += 1
            if positive > 5:
                dt = min(dt * 1.1, dt_max)
                alpha *= 0.99
        else:
            velocity.fill(0.0)
            positive = 0
            alpha = 0.1
            if it > 0:
                dt *= 0.5
        velocity += dt * force
        step = dt * velocity
        largest = np.max(np.linalg.norm(step, axis=1))
        if largest > 0.1:
            step *= 0.1 / largest
        x = x + step
        trajectory.append(x.copy())
        force = force_fn(x)
    return x, False, max_iterations, trajectory
