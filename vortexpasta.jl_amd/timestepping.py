"""Time stepping of the filaments with the SSPRK33 scheme.

Mirrors ``Timestepping.step!`` with ``SSPRK33`` (``src/Timestepping/Timestepping.jl:1038-1076``,
``timesteppers/explicit/SSPRK33.jl:17-39``) and the filament advection of
``advect.jl:20-72``, for the Biot–Savart velocity alone (no forcing,
reconnections or adaptivity):

    v₁ = v(s);  s₂ = s + Δt v₁;  v₂ = v(s₂)
    s₃ = s + Δt/4 (v₁ + v₂);  v₃ = v(s₃)
    s ← s + Δt (v₁ + v₂ + 4 v₃)/6

Within a step the intermediate filaments keep the knots of s (only the spline
coefficients are solved again); after the step the knots are recomputed from
the new nodes (chordal parametrisation).

Two drivers: :func:`ssprk33_step` (host: the filaments live on the host, each
stage calls ``velocity``) and :class:`DeviceStepper` (``vpbs_advance``: nodes,
knots and coefficients stay on the GPU between stages and steps).

There is no adaptive time step: Δt must resolve the fastest Kelvin waves at
the node spacing δ, ω ≈ Γ k²/(4π) ln(1/(k a)) with k = π/δ (the scheme's
stability region ends near ω·Δt ≈ 2.5).
"""

from __future__ import annotations

from typing import Callable

import numpy as np

from .biotsavart import BiotSavartCache, velocity_on_nodes
from .filaments import FilamentBatch

__all__ = ["ssprk33_step", "ssprk33_update", "DeviceStepper"]


def ssprk33_update(x: np.ndarray, rhs: Callable[[np.ndarray, bool], np.ndarray], dt: float) -> np.ndarray:
    """One SSPRK33 step of dx/dt = rhs(x) (SSPRK33.jl:17-39).

    ``rhs(x, keep_knots)`` is told whether x is an intermediate stage (knots kept).
    """
    v1 = rhs(x, False)
    v2 = rhs(x + dt * v1, True)
    vs = v1 + v2
    v3 = rhs(x + (dt / 4) * vs, True)
    v = (vs + 4 * v3) / 6
    return x + dt * v


def ssprk33_step(batch: FilamentBatch, dt: float, velocity: Callable[[FilamentBatch], np.ndarray]) -> FilamentBatch:
    """Advance all filaments by one step; ``velocity(batch)`` returns the (Np, 3) node velocities."""
    X0 = batch.X

    def rhs(x, stage):
        b = batch if x is X0 else batch.with_nodes(x, keep_knots=True)
        return velocity(b)

    X1 = ssprk33_update(X0, rhs, dt)
    return batch.with_nodes(X1, keep_knots=False)


def cache_velocity(cache: BiotSavartCache):
    """``velocity(batch)`` through the GPU library (compute_on_nodes_)."""

    def velocity(b: FilamentBatch) -> np.ndarray:
        return np.concatenate(velocity_on_nodes(cache, b), axis=0)

    return velocity


class DeviceStepper:
    """SSPRK33 on device-resident filaments (``vpbs_advance``).

    The filaments are uploaded once (``vpbs_set_filaments``); each step then
    runs entirely on the GPU: point data, velocity, stage advection, spline
    coefficients (and chordal knots after the step).
    """

    def __init__(self, cache: BiotSavartCache, batch: FilamentBatch):
        self.cache = cache
        self.batch = batch
        cache.ctx.set_filaments(batch, cache.params)

    def advance(self, dt: float, nsteps: int = 1):
        self.cache.ctx.advance(dt, nsteps)

    def nodes(self) -> np.ndarray:
        """Current nodes (Np, 3), not folded."""
        return self.cache.ctx.get_filaments()["X"]

    def get_batch(self) -> FilamentBatch:
        st = self.cache.ctx.get_filaments()
        return self.batch.with_nodes(st["X"], keep_knots=False)
