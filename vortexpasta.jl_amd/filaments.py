"""Closed vortex filaments discretised by periodic quintic splines.

Host-side mirror of the parts of the reference's ``Filaments`` module that the
Biot–Savart path reads:

* nodes at ``τ_i = (2i - 1)/(2N)`` of a parametric curve
  (``src/Filaments/closed_filament.jl:183-187``), with the end-to-end offset
  ``S(1) - S(0)`` of unclosed periodic lines (``closed_filament.jl:169-173``);
* chordal knots ``t_{i+1} = t_i + |X_{i+1} - X_i|`` (``Filaments.jl:695-708``,
  ``parametrisations.jl`` ``ChordalParametrisation``);
* the periodic quintic (order-6) B-spline interpolant through the nodes
  (``spline/spline.jl:96-139,264-319``, ``spline/quintic.jl``), evaluated with
  the Cox–de Boor recursion; unclosed lines are periodised by
  ``Y = X - (t/T)·Xoffset`` as in ``spline.jl:163-193``;
* predefined curves ``Ring``, ``TrefoilKnot``, ``PeriodicLine`` and
  ``define_curve`` (``src/PredefinedCurves/PredefinedCurves.jl:205-224``,
  ``ring.jl:100-108``, ``trefoil.jl:24-33``, ``periodic_line.jl:110-120``).

The interpolant is unique given nodes and knots, so this module reproduces the
reference's curve up to roundoff; the implementation (a sparse block-cyclic
solve over all filaments at once and vectorised basis evaluation) is this
package's own.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

__all__ = [
    "Ring",
    "TrefoilKnot",
    "PeriodicLine",
    "define_curve",
    "ClosedFilament",
    "FilamentBatch",
    "fold_coordinates_periodic",
    "init",
]

ORDER = 6  # quintic spline (SplineMethod{6}, spline.jl:60)


# --------------------------------------------------------------------------- #
# Parametric curves (period 1 in t)


class Ring:
    """Unit circle in the XY plane: S(t) = (cos 2πt, sin 2πt, 0) (``ring.jl:100-108``)."""

    def __call__(self, t):
        u = 2 * np.asarray(t, dtype=np.float64)
        return np.stack([np.cos(np.pi * u), np.sin(np.pi * u), np.zeros_like(u)], axis=-1)


class TrefoilKnot:
    """Trefoil knot (``trefoil.jl:24-33``)."""

    def __call__(self, t):
        u = 2 * np.asarray(t, dtype=np.float64)
        return np.stack(
            [
                np.sin(np.pi * u) + 2 * np.sin(2 * np.pi * u),
                np.cos(np.pi * u) - 2 * np.cos(2 * np.pi * u),
                -np.sin(3 * np.pi * u),
            ],
            axis=-1,
        )


class PeriodicLine:
    """Unclosed line along z with periodic perturbations (``periodic_line.jl:110-120``).

    ``x``, ``y`` are real and ``r`` complex functions of t with period 1.
    """

    def __init__(self, x: Optional[Callable] = None, y: Optional[Callable] = None, r: Optional[Callable] = None):
        self.x = x
        self.y = y
        self.r = r

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        x = self.x(t) if self.x is not None else np.zeros_like(t)
        y = self.y(t) if self.y is not None else np.zeros_like(t)
        if self.r is not None:
            rt = self.r(t)
            x = x + np.real(rt)
            y = y + np.imag(rt)
        return np.stack([x, y, t - 0.5], axis=-1)


def define_curve(curve, *, translate=0.0, scale=1.0, rotate=None, orientation: int = 1):
    """S(t) = translate + rotate·scale·S_base(τ) (``PredefinedCurves.jl:205-224``)."""
    translate = np.broadcast_to(np.asarray(translate, dtype=np.float64), (3,)).copy()
    if np.isscalar(scale):
        A = float(scale) * np.eye(3)
    else:
        scale = np.asarray(scale, dtype=np.float64)
        A = np.diag(scale) if scale.ndim == 1 else scale
    if rotate is not None:
        A = np.asarray(rotate, dtype=np.float64) @ A

    def S(t):
        t = np.asarray(t, dtype=np.float64)
        tau = t if orientation > 0 else 1 - t
        x = curve(tau)
        return translate + x @ A.T

    return S


# --------------------------------------------------------------------------- #
# Folding


def fold_coordinates_periodic(x: np.ndarray, L: float) -> np.ndarray:
    """Fold into [0, L) with the reference's loop order (``min_distance.jl:5-20``).

    The ``x < 0`` loop runs first so that a tiny negative value ends up at 0,
    never at L.
    """
    x = np.array(x, dtype=np.float64, copy=True)
    neg = x < 0
    while neg.any():
        x[neg] += L
        neg = x < 0
    big = x >= L
    while big.any():
        x[big] -= L
        big = x >= L
    return x


# --------------------------------------------------------------------------- #
# B-spline machinery (vectorised over evaluation points)


def _knot_ext(ts: np.ndarray, off: np.ndarray, N: np.ndarray, T: np.ndarray, j: np.ndarray) -> np.ndarray:
    """Periodically extended knot t_j = t_{j mod N} + ⌊j/N⌋·T (per-filament arrays)."""
    q, r = np.divmod(j, N)
    return ts[off + r] + q * T


def _basis(k: int, t, i, knot):
    """Non-zero order-k B-splines at t ∈ [t_i, t_{i+1}).

    Returns an array B of shape (npts, k) with B[:, r] = B_{i-k+1+r, k}(t)
    (Cox–de Boor triangle; ``spline.jl:199-253`` evaluates the same basis).
    """
    npts = t.shape[0]
    Nb = np.zeros((npts, k))
    Nb[:, 0] = 1.0
    left = [None] * k
    right = [None] * k
    for j in range(1, k):
        left[j] = t - knot(i + 1 - j)
        right[j] = knot(i + j) - t
        saved = np.zeros(npts)
        for r in range(j):
            temp = Nb[:, r] / (right[r + 1] + left[j - r])
            Nb[:, r] = saved + right[r + 1] * temp
            saved = left[j - r] * temp
        Nb[:, j] = saved
    return Nb


@dataclass
class ClosedFilament:
    """A closed (or periodically unclosed) filament with N ≥ 5 nodes."""

    nodes: np.ndarray  # (N, 3)
    offset: np.ndarray  # (3,) end-to-end offset f[N] - f[0]

    def __post_init__(self):
        self.nodes = np.ascontiguousarray(self.nodes, dtype=np.float64)
        self.offset = np.asarray(self.offset, dtype=np.float64).reshape(3)
        if self.nodes.ndim != 2 or self.nodes.shape[1] != 3:
            raise ValueError("nodes must have shape (N, 3)")
        if len(self.nodes) < 5:  # spline.jl:64
            raise ValueError("quintic splines need at least 5 nodes")

    def __len__(self) -> int:
        return len(self.nodes)

    def end_to_end_offset(self) -> np.ndarray:
        return self.offset


def init(S: Callable, N: int) -> ClosedFilament:
    """Filament with N nodes at τ_i = (2i-1)/(2N) of the curve S (``closed_filament.jl:183-187``)."""
    tau = (2 * np.arange(1, N + 1) - 1) / (2 * N)
    nodes = S(tau)
    offset = S(np.array([1.0]))[0] - S(np.array([0.0]))[0]
    # Closed curves give an offset of rounding size: treat it as exactly closed.
    if np.all(np.abs(offset) < 1e-12 * max(1.0, np.abs(nodes).max())):
        offset = np.zeros(3)
    return ClosedFilament(nodes, offset)


class FilamentBatch:
    """All filaments of a system, with spline coefficients solved together.

    Attributes are flat arrays over filaments: node offsets ``off`` (start of
    each filament in the linear node index), counts ``N``, knot periods ``T``
    and offsets ``Xoff``.
    """

    def __init__(self, fs: Sequence[ClosedFilament], knots=None):
        """``knots``: optional ``(ts, T)`` to keep instead of recomputing chordal knots
        (the reference keeps the knots within Runge–Kutta stages, ``advect.jl:47-61``)."""
        if len(fs) == 0:
            raise ValueError("empty filament list")
        self.fs = list(fs)
        Ns = np.array([len(f) for f in fs], dtype=np.int64)
        self.Nf = len(fs)
        self.Ns = Ns
        self.offs = np.concatenate([[0], np.cumsum(Ns)[:-1]]).astype(np.int64)
        self.Np = int(Ns.sum())
        self.X = np.concatenate([f.nodes for f in fs], axis=0)  # (Np, 3)
        self.Xoff = np.stack([f.offset for f in fs], axis=0)  # (Nf, 3)
        fid = np.repeat(np.arange(self.Nf), Ns)
        self.fid = fid
        self.local = np.arange(self.Np) - self.offs[fid]
        if knots is None:
            self._compute_knots()
        else:
            self.ts = np.array(knots[0], dtype=np.float64)
            self.T = np.array(knots[1], dtype=np.float64)
            if self.ts.shape != (self.Np,) or self.T.shape != (self.Nf,):
                raise ValueError("knots do not match the filaments")
        self._solve_coefficients()

    def with_nodes(self, X: np.ndarray, keep_knots: bool) -> "FilamentBatch":
        """Same filaments with node positions X (Np, 3); knots kept or recomputed (chordal)."""
        fs = [ClosedFilament(X[a : a + n], f.offset) for f, a, n in zip(self.fs, self.offs, self.Ns)]
        return FilamentBatch(fs, knots=(self.ts, self.T) if keep_knots else None)

    # per-node filament properties
    def _fN(self, fid):
        return self.Ns[fid]

    def _compute_knots(self):
        X, fid, offs, Ns = self.X, self.fid, self.offs, self.Ns
        nxt = np.where(self.local == Ns[fid] - 1, offs[fid], np.arange(self.Np) + 1)
        Xn = X[nxt] + np.where((self.local == Ns[fid] - 1)[:, None], self.Xoff[fid], 0.0)
        dt = np.sqrt(np.sum((Xn - X) ** 2, axis=1))  # chordal increments
        ts = np.empty(self.Np)
        T = np.empty(self.Nf)
        for f in range(self.Nf):  # cumulative sum per filament, t_0 = 0
            a, n = offs[f], Ns[f]
            c = np.cumsum(dt[a : a + n])
            ts[a] = 0.0
            ts[a + 1 : a + n] = c[:-1]
            T[f] = c[-1]
        self.ts = ts
        self.T = T

    def knot(self, fid, j):
        return _knot_ext(self.ts, self.offs[fid], self.Ns[fid], self.T[fid], j)

    def _solve_coefficients(self):
        """Solve the block-cyclic interpolation system A·c = Y (5 non-zeros per row)."""
        fid, i = self.fid, self.local
        t = self.ts
        knot = lambda j: self.knot(fid, j)  # noqa: E731
        B = _basis(ORDER, t, i, knot)  # B[:, r] = B_{i-5+r}(t_i), B[:, 5] = 0
        rows = np.repeat(np.arange(self.Np), ORDER - 1)
        cols_local = (i[:, None] - (ORDER - 1) + np.arange(ORDER - 1)[None, :]) % self.Ns[fid][:, None]
        cols = (self.offs[fid][:, None] + cols_local).ravel()
        vals = B[:, : ORDER - 1].ravel()
        A = sp.csc_matrix((vals, (rows, cols)), shape=(self.Np, self.Np))
        Y = self.X - (t / self.T[fid])[:, None] * self.Xoff[fid]  # periodised coordinates
        lu = spla.splu(A)
        c = np.empty_like(Y)
        for d in range(3):
            c[:, d] = lu.solve(np.ascontiguousarray(Y[:, d]))
        self.c = c
        # derivative coefficients: d_j = 5(c_j - c_{j-1})/(t_{j+5} - t_j), e_j = 4(d_j - d_{j-1})/(t_{j+4} - t_j)
        jprev = self.offs[fid] + (i - 1) % self.Ns[fid]
        den5 = knot(i + 5) - knot(i)
        self.d = 5 * (c - c[jprev]) / den5[:, None]
        den4 = knot(i + 4) - knot(i)
        self.e = 4 * (self.d - self.d[jprev]) / den4[:, None]

    def evaluate(self, fid: np.ndarray, seg: np.ndarray, zeta: np.ndarray, deriv: int = 0) -> np.ndarray:
        """Evaluate S^(deriv) at t = (1-ζ)t_i + ζt_{i+1} on segment ``seg`` of filament ``fid``.

        Mirrors ``spline.jl:116-122`` (local parametrisation ζ ∈ [0, 1]).
        """
        fid = np.asarray(fid, dtype=np.int64)
        seg = np.asarray(seg, dtype=np.int64)
        zeta = np.asarray(zeta, dtype=np.float64)
        knot = lambda j: self.knot(fid, j)  # noqa: E731
        ti = knot(seg)
        ti1 = knot(seg + 1)
        t = (1 - zeta) * ti + zeta * ti1
        k = ORDER - deriv
        coefs = (self.c, self.d, self.e)[deriv]
        Bk = _basis(k, t, seg, knot)
        N = self.Ns[fid]
        off = self.offs[fid]
        out = np.zeros((len(t), 3))
        for r in range(k):
            idx = off + (seg - k + 1 + r) % N
            out += Bk[:, r : r + 1] * coefs[idx]
        # undo periodisation of unclosed lines (spline.jl:185-193)
        if deriv == 0:
            out += (t / self.T[fid])[:, None] * self.Xoff[fid]
        elif deriv == 1:
            out += self.Xoff[fid] / self.T[fid][:, None]
        return out

    def segment_dt(self) -> np.ndarray:
        """Δt_i = t_{i+1} - t_i for every node's right segment."""
        return self.knot(self.fid, self.local + 1) - self.ts

    def filaments_from_linear(self, v: np.ndarray) -> List[np.ndarray]:
        """Split a linear (Np, 3) array into one array per filament."""
        return [v[a : a + n] for a, n in zip(self.offs, self.Ns)]
