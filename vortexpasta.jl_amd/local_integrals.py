"""Near-singular local integrals for ``lia_segment_fraction`` (host side).

Mirrors ``add_local_integrals!`` (``src/BiotSavart/shortrange/local_integrals.jl:209-305``),
which the reference also runs on the host, synchronously, next to the device
work (``BiotSavart.jl:570-580``). When ``lia_segment_fraction = γ`` the LIA term
covers only the fraction γ of each segment next to a node (the δ± of the point
data are computed over those sub-segments); the rest of the two segments —
ζ ∈ [0, 1-γ] of segment i-1 and ζ ∈ [γ, 1] of segment i
(``nonlia_integration_limits``, ``local_integrals.jl:103-104``) — is integrated
with the full Biot–Savart kernel using ``quadrature_near_singularity``.
"""

from __future__ import annotations

import numpy as np

from .filaments import fold_coordinates_periodic


def add_local_integrals(v, psi, batch, params) -> None:
    """Add the non-LIA part of the local segments to v and/or psi ((3, Np) arrays, in place).

    ``batch`` is the FilamentBatch the point data was built from.
    """
    g = params.lia_segment_fraction
    if g is None:
        return
    b = batch
    zs, ws = params.quadrature_near_singularity.rule()
    Ls = np.asarray(params.Ls, dtype=np.float64)[:, None]
    x = np.stack([fold_coordinates_periodic(b.X[:, d], Ls[d, 0]) for d in range(3)])  # folded nodes (3, Np)
    N = b.Ns[b.fid]
    pref = params.Γ / (4 * np.pi)
    acc_v = np.zeros_like(x)
    acc_p = np.zeros_like(x)
    for seg, (lo, hi) in (((b.local - 1) % N, (0.0, 1.0 - g)), (b.local, (g, 1.0))):
        delta = hi - lo
        dt = (b.knot(b.fid, seg + 1) - b.knot(b.fid, seg)) * delta  # local_integrals.jl:230
        for z, w in zip(zs, ws):
            y = np.full(b.Np, lo + delta * z)
            s = b.evaluate(b.fid, seg, y, 0).T
            q = (w * dt)[None, :] * b.evaluate(b.fid, seg, y, 1).T
            s = np.stack([fold_coordinates_periodic(s[d], Ls[d, 0]) for d in range(3)])
            r = x - s
            r = np.where(r >= Ls / 2, r - Ls, r)
            r = np.where(r < -Ls / 2, r + Ls, r)
            rinv = 1.0 / np.sqrt(np.sum(r * r, axis=0))
            if v is not None:
                acc_v += np.cross(q.T, r.T).T * rinv ** 3
            if psi is not None:
                acc_p += q * rinv
    if v is not None:
        v += pref * acc_v
    if psi is not None:
        psi += pref * acc_p
