"""Forcing and dissipation terms built on the long-range Fourier field.

Mirrors ``Forcing.FourierBandForcingBS`` (``src/Forcing/fourier_band_bs.jl``)
and ``Forcing.SmallScaleDissipationBS``
(``src/Forcing/dissipation_bs_small_scales.jl``): both read the unsmoothed
Biot–Savart velocity in Fourier space that the last evaluation kept. The
filtered velocity at the nodes is computed on the GPU from that field
(``vpbs_interpolate_band``: band or low-pass filter, inverse transforms and
interpolation); the per-node algebra below is O(Np).

Differences of form: the reference evaluates the band field by a direct sum
over the band's wavevectors (one half of them, real part); here it is a type-2
NUFFT of the band-limited field, equal to that sum to the NUFFT tolerance. The
reference's half-space real part is half of the full band field:
``Re Σ_{q ∈ half band} v̂ e^{ik·s} = ½ Σ_{k ∈ band} v̂ e^{ik·s}`` for k ≠ 0.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .biotsavart import BiotSavartCache
from .filaments import FilamentBatch

__all__ = ["FourierBandForcingBS", "SmallScaleDissipationBS", "forcing_velocity", "dissipation_velocity"]


@dataclass(frozen=True)
class FourierBandForcingBS:
    """Forcing in the band kmin ≤ |k| ≤ kmax (``fourier_band_bs.jl:130-150``): pass α or ε_target, not both.

    ``α_prime`` is the reference's ``α′`` (a prime is not a valid Python identifier character).
    """

    kmin: float
    kmax: float
    α: float = 0.0
    ε_target: float = 0.0
    α_prime: float = 0.0
    modify_length: bool = True

    def __post_init__(self):
        if (self.α == 0) + (self.ε_target == 0) != 1:
            raise ValueError("one should pass either α or ε_target, but not both")
        if not 0 <= self.kmin <= self.kmax:
            raise ValueError("expected 0 < kmin ≤ kmax")


@dataclass(frozen=True)
class SmallScaleDissipationBS:
    """Dissipation at k > kdiss (``dissipation_bs_small_scales.jl:57-67``): pass α or ε_target, not both."""

    kdiss: float
    α: float = 0.0
    ε_target: float = 0.0

    def __post_init__(self):
        if (self.α == 0) + (self.ε_target == 0) != 1:
            raise ValueError("one should pass either α or ε_target, but not both")


def _node_geometry(batch: FilamentBatch):
    """Unit tangent s′, parametric s″ at nodes and the chordal lengths of the two local segments."""
    zeros = np.zeros(batch.Np)
    st = batch.evaluate(batch.fid, batch.local, zeros, deriv=1)
    stt = batch.evaluate(batch.fid, batch.local, zeros, deriv=2)
    sp = st / np.linalg.norm(st, axis=1)[:, None]
    N = batch.Ns[batch.fid]
    off = batch.offs[batch.fid]
    nxt = np.where(batch.local == N - 1, off, np.arange(batch.Np) + 1)
    Xn = batch.X[nxt] + np.where((batch.local == N - 1)[:, None], batch.Xoff[batch.fid], 0.0)
    seg = np.linalg.norm(Xn - batch.X, axis=1)  # chordal length of the right segment
    prv = np.where(batch.local == 0, off + N - 1, np.arange(batch.Np) - 1)
    return st, stt, sp, seg, seg[prv]


def forcing_velocity(forcing: FourierBandForcingBS, cache: BiotSavartCache, fs: Sequence) -> List[np.ndarray]:
    """Forcing velocity on every node (``fourier_band_bs.jl:293-420``), one (N_i, 3) array per filament.

    Uses the Fourier field of the last evaluation of ``fs`` with ``cache``.
    """
    batch = cache.batch if cache.batch is not None and cache.batch.Np == sum(len(f) for f in fs) else FilamentBatch(fs)
    if cache.ctx.uhat_quantity() != "velocity":
        raise RuntimeError("the forcing needs the unsmoothed velocity in Fourier space (evaluate the velocity first)")
    band = cache.ctx.interpolate_band(forcing.kmin, forcing.kmax).T  # (Np, 3), full band field
    vs = 0.5 * band  # the reference's half-space real part
    st, stt, sp, seg, seg_prev = _node_geometry(batch)
    dxi = seg_prev + seg  # length of the two local segments (fourier_band_bs.jl:232-243)
    α, α_prime = forcing.α, forcing.α_prime
    if forcing.modify_length:
        vf = np.cross(sp, vs)
        eps_local = np.sum(vf * vf, axis=1) * dxi
    else:
        s2 = np.cross(sp, np.cross(stt, sp))  # curvature direction (up to a factor)
        n2 = np.sum(s2 * s2, axis=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            c = np.where(n2 == 0, 0.0, np.sum(s2 * vs, axis=1) / np.where(n2 == 0, 1.0, n2))
        vf = c[:, None] * np.cross(sp, s2)
        with np.errstate(divide="ignore", invalid="ignore"):
            eps_local = np.where(n2 == 0, 0.0, np.sum(s2 * vs, axis=1) ** 2 / np.where(n2 == 0, 1.0, n2)) * dxi
    if α != 0:
        vf = α * vf - α_prime * np.cross(sp, vf)
    else:
        prefactor = cache.params.Γ / float(np.prod(cache.params.Ls))
        eps_total = prefactor * eps_local.sum()
        if eps_total == 0:
            vf = np.zeros_like(vf)
        else:
            a = forcing.ε_target / eps_total
            vf = a * vf if α_prime == 0 else a * vf - α_prime * np.cross(sp, a * vf)
    return batch.filaments_from_linear(vf)


def dissipation_velocity(dissipation: SmallScaleDissipationBS, cache: BiotSavartCache, fs: Sequence,
                         vs_all: Sequence[np.ndarray]) -> List[np.ndarray]:
    """Dissipation velocity on every node (``dissipation_bs_small_scales.jl:131-190``).

    ``vs_all``: the Biot–Savart velocity on the nodes from the same evaluation.
    """
    batch = cache.batch if cache.batch is not None and cache.batch.Np == sum(len(f) for f in fs) else FilamentBatch(fs)
    Ls, Ns = cache.params.Ls, cache.params.Ns
    for d in range(3):  # dissipation_bs_small_scales.jl:86-97
        kmax = 2 * np.pi / Ls[d] * ((Ns[d] // 2 - 1) if d == 0 else (Ns[d] - 1) // 2)
        if dissipation.kdiss >= kmax:
            raise ValueError(f"SmallScaleDissipationBS: dissipative wavenumber (kdiss = {dissipation.kdiss}) is too "
                             f"large compared to size of Fourier grid (kmax = {kmax})")
    if cache.ctx.uhat_quantity() != "velocity":
        raise RuntimeError("the dissipation needs the unsmoothed velocity in Fourier space")
    vlow = cache.ctx.interpolate_band(0.0, dissipation.kdiss).T  # k² ≤ kdiss²
    vs = np.concatenate([np.asarray(v) for v in vs_all], axis=0)
    st, stt, sp, seg, seg_prev = _node_geometry(batch)
    vdiss = np.cross(vs - vlow, sp)  # = -s′ × v>
    if dissipation.α != 0:
        vdiss = dissipation.α * vdiss
    else:
        t_prev = batch.knot(batch.fid, batch.local - 1)
        t_next = batch.knot(batch.fid, batch.local + 1)
        dxi = (t_next - t_prev) / 2 * np.linalg.norm(st, axis=1)
        eps_total = cache.params.Γ / float(np.prod(Ls)) * np.sum(np.sum(vdiss * vdiss, axis=1) * dxi)
        if eps_total != 0:
            vdiss = dissipation.ε_target / eps_total * vdiss
    return batch.filaments_from_linear(vdiss)
