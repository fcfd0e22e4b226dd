"""Linearised point data: nodes, quadrature points and vector charges.

Mirrors ``PointData`` (``src/BiotSavart/pointdata.jl:21-39``) and
``add_point_charges!`` (``src/BiotSavart/pointdata_impl.jl:121-188``):

* ``nodes``: filament nodes folded into [0, L)³ (``pointdata_impl.jl:161``);
* ``node_idx_prev``: linear index of the node to the left, wrapping around each
  closed filament (``:162-167``), 0-based here;
* ``sp``, ``spp``: s′ and s″ at nodes (``:168-169``);
* ``dlm``, ``dlp``: sub-segment lengths δ⁻, δ⁺ from the same quadrature rule
  (``:133-143,170-172``), over ``lia_integration_limits`` when
  ``lia_segment_fraction`` is set (``local_integrals.jl:100-101``);
* ``points``, ``charges``: for segment i and quadrature node (ζ, w),
  s(ζ) folded and w·Δt·s′(ζ) (``:173-183``). Points of segment i occupy
  indices ``Nq*i .. Nq*i + Nq - 1`` (0-based), which the local-segment tests
  of the pair kernel rely on (``shortrange.jl:150-153``).

All arrays are float64 structure-of-arrays ``(3, n)``, the layout the C ABI
takes.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .filaments import FilamentBatch, fold_coordinates_periodic
from .params import NoQuadrature

__all__ = ["PointData", "add_point_charges"]


@dataclass
class PointData:
    nodes: np.ndarray  # (3, Np) folded
    node_idx_prev: np.ndarray  # (Np,) int32
    sp: np.ndarray  # (3, Np)
    spp: np.ndarray  # (3, Np)
    dlm: np.ndarray  # (Np,)
    dlp: np.ndarray  # (Np,)
    points: np.ndarray  # (3, Nq) folded
    charges: np.ndarray  # (3, Nq)
    nquad: int
    batch: FilamentBatch

    @property
    def Np(self) -> int:
        return self.nodes.shape[1]

    @property
    def Nq(self) -> int:
        return self.points.shape[1]


def _fold3(x: np.ndarray, Ls) -> np.ndarray:
    return np.stack([fold_coordinates_periodic(x[:, d], Ls[d]) for d in range(3)], axis=0)


def _lia_limits(gamma):
    if gamma is None:
        return (0.0, 1.0), (0.0, 1.0)
    return (1.0 - gamma, 1.0), (0.0, gamma)  # local_integrals.jl:101


def add_point_charges(fs, params, batch: FilamentBatch = None) -> PointData:
    """Build point data for all filaments (``pointdata_impl.jl:121-131``)."""
    if batch is None:
        batch = FilamentBatch(fs)
    Ls = params.Ls
    quad = params.quadrature
    Np = batch.Np
    fid, loc, N, off = batch.fid, batch.local, batch.Ns[batch.fid], batch.offs[batch.fid]

    nodes = _fold3(batch.X, Ls)
    idx_prev = np.where(loc == 0, off + N - 1, np.arange(Np) - 1).astype(np.int32)

    zeros = np.zeros(Np)
    sp = batch.evaluate(fid, loc, zeros, deriv=1)
    spp = batch.evaluate(fid, loc, zeros, deriv=2)

    dts = batch.segment_dt()  # Δt of the right segment of each node
    if isinstance(quad, NoQuadrature):
        # Straight segments, midpoint (pointdata_impl.jl:190-219)
        gamma = 1.0 if params.lia_segment_fraction is None else params.lia_segment_fraction
        nxt = np.where(loc == N - 1, off, np.arange(Np) + 1)
        Xn = batch.X[nxt] + np.where((loc == N - 1)[:, None], batch.Xoff[fid], 0.0)
        Xp = batch.X[idx_prev] - np.where((loc == 0)[:, None], batch.Xoff[fid], 0.0)
        dlm = gamma * np.linalg.norm(batch.X - Xp, axis=1)
        dlp = gamma * np.linalg.norm(Xn - batch.X, axis=1)
        pts = (batch.X + Xn) / 2
        chg = Xn - batch.X
        return PointData(
            nodes, idx_prev, sp.T.copy(), spp.T.copy(), dlm, dlp, _fold3(pts, Ls), chg.T.copy(), 1, batch
        )

    zs, ws = quad.rule()
    nq = len(zs)
    lims_m, lims_p = _lia_limits(params.lia_segment_fraction)

    def seg_length(seg_local, lims):
        a, b = lims
        delta = b - a
        segl = seg_local % N
        acc = np.zeros(Np)
        dt = batch.knot(fid, segl + 1) - batch.knot(fid, segl)
        for z, w in zip(zs, ws):
            d1 = batch.evaluate(fid, segl, np.full(Np, a + delta * z), deriv=1)
            acc = acc + w * np.sqrt(np.sum(d1 * d1, axis=1))
        return dt * (delta * acc)  # integrate.jl:40-61, Quadratures.jl:131-145

    dlm = seg_length(loc - 1, lims_m)
    dlp = seg_length(loc, lims_p)

    # Quadrature points and charges, segment-major (Nq*i + q).
    seg = np.repeat(loc, nq)
    fq = np.repeat(fid, nq)
    zq = np.tile(zs, Np)
    wq = np.tile(ws, Np)
    s = batch.evaluate(fq, seg, zq, deriv=0)
    s1 = batch.evaluate(fq, seg, zq, deriv=1)
    q = (wq * np.repeat(dts, nq))[:, None] * s1
    return PointData(
        nodes, idx_prev, sp.T.copy(), spp.T.copy(), dlm, dlp, _fold3(s, Ls), q.T.copy(), nq, batch
    )
