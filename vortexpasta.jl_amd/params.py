"""Biot–Savart parameters: Ewald splitting, quadrature and backend selection.

Host-side mirror of the reference's parameter layer:

* ``GaussianSplitting`` and its derivation rules follow
  ``src/BiotSavart/splitting/gaussian.jl:109-200`` (rcut = β/α, kmax = 2αβ,
  rtol = 10^-1.3 · exp(-β²)).
* ``ParamsBiotSavart`` follows ``src/BiotSavart/params.jl:326-376`` (keyword
  names, defaults and validity checks of ``params.jl:99-107``).
* ``kmax_to_gridsize`` / ``maximum_wavenumber`` follow ``params.jl:156-166``.
* ``GaussLegendre`` follows ``src/Quadratures/Quadratures.jl:52-56,147-169``
  (nodes and weights mapped to [0, 1]).

The two backend types are the drop-in replacements for the reference's
``CellListsBackend`` (``shortrange/backends/cell_lists.jl:64-93``) and
``NonuniformFFTsBackend`` (``longrange/backends/nonuniformffts.jl:97-140``):
selecting them is the only change a user makes.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np

__all__ = [
    "GaussLegendre",
    "NoQuadrature",
    "GaussianSplitting",
    "HIPShortRangeBackend",
    "HIPLongRangeBackend",
    "ParamsBiotSavart",
    "kmax_to_gridsize",
    "maximum_wavenumber",
    "max_cutoff_distance_celllists",
]

_C_RTOL = 10.0 ** (-1.3)  # empirical factor of rtol(β) ≈ C·exp(-β²) (gaussian.jl:139)


# --------------------------------------------------------------------------- #
# Quadratures


@dataclass(frozen=True)
class GaussLegendre:
    """N-point Gauss–Legendre rule on [0, 1] (``Quadratures.jl:74-76,147-156``)."""

    n: int

    def __post_init__(self):
        if self.n < 1:
            raise ValueError("GaussLegendre needs at least one point")

    def __len__(self) -> int:
        return self.n

    def rule(self) -> Tuple[np.ndarray, np.ndarray]:
        xs, ws = np.polynomial.legendre.leggauss(self.n)
        # [-1, 1] -> [0, 1] (Quadratures.jl:50-55)
        return (xs + 1.0) / 2.0, ws / 2.0


@dataclass(frozen=True)
class NoQuadrature:
    """Midpoint rule on straight segments (``Quadratures.jl:78-88``).

    Point data generation supports it (``pointdata_impl.jl:190-219``); it is
    here so that parameter objects can be built as in the reference.
    """

    def __len__(self) -> int:
        return 1

    def rule(self):
        return np.array([0.5]), np.array([1.0])


# --------------------------------------------------------------------------- #
# Ewald splitting


def kmax_to_gridsize(kmax: float, L: float) -> int:
    """Grid size for a wanted kmax (``params.jl:156-159``)."""
    m = math.ceil(kmax * L / (2 * math.pi))
    return 2 * m + 2


def maximum_wavenumber(Ns: Sequence[int], Ls: Sequence[float]) -> float:
    """kmax = min_d 2π·((N_d - 1) ÷ 2)/L_d (``params.jl:161-166``)."""
    return min(2 * math.pi * ((N - 1) // 2) / L for N, L in zip(Ns, Ls))


def _ns_from_kmax(Ls, kmax):
    return tuple(kmax_to_gridsize(kmax, L) for L in Ls)


@dataclass(frozen=True)
class GaussianSplitting:
    """Gaussian Ewald splitting (``splitting/gaussian.jl:109-200``).

    Build it with ``GaussianSplitting.create(Ls=..., β=..., Ns=...)`` (or any of
    the argument combinations accepted by the reference). The fields hold the
    derived values.
    """

    Ls: Tuple[float, float, float]
    α: float
    rcut: float
    Ns: Tuple[int, int, int]
    rtol: float

    @staticmethod
    def create(*, Ls, β=None, α=None, rcut=None, Ns=None, rtol=None) -> "GaussianSplitting":
        Ls = _as_periods(Ls)
        Ns = None if Ns is None else _as_dims(Ns)
        given = (β is not None, α is not None, rcut is not None, Ns is not None, rtol is not None)
        if given == (True, True, False, False, False):  # gaussian.jl:133-142
            rcut = β / α
            Ns = _ns_from_kmax(Ls, 2 * α * β)
            rtol = _C_RTOL * math.exp(-β * β)
        elif given == (True, False, True, False, False):  # gaussian.jl:144-153
            α = β / rcut
            Ns = _ns_from_kmax(Ls, 2 * α * β)
            rtol = _C_RTOL * math.exp(-β * β)
        elif given == (True, False, False, True, False):  # gaussian.jl:155-162
            kmax = maximum_wavenumber(Ns, Ls)
            α = kmax / (2 * β)
            rcut = β / α
            rtol = _C_RTOL * math.exp(-β * β)
        elif given == (False, True, True, True, False):  # gaussian.jl:164-169
            b = α * rcut
            rtol = _C_RTOL * math.exp(-b * b)
        elif given == (False, True, False, False, True):  # gaussian.jl:171-180
            b = math.sqrt(-math.log(rtol / _C_RTOL))
            rcut = b / α
            Ns = _ns_from_kmax(Ls, 2 * α * b)
        elif given == (False, False, True, False, True):  # gaussian.jl:182-191
            b = math.sqrt(-math.log(rtol / _C_RTOL))
            α = b / rcut
            Ns = _ns_from_kmax(Ls, 2 * α * b)
        elif given == (False, False, False, True, True):  # gaussian.jl:193-200
            b = math.sqrt(-math.log(rtol / _C_RTOL))
            kmax = maximum_wavenumber(Ns, Ls)
            α = kmax / (2 * b)
            rcut = b / α
        else:
            raise ValueError(
                "GaussianSplitting: pass β or rtol together with one of α, rcut, Ns "
                "(or all of α, rcut, Ns)"
            )
        return GaussianSplitting(Ls, float(α), float(rcut), tuple(int(n) for n in Ns), float(rtol))

    @property
    def kmax(self) -> float:
        return maximum_wavenumber(self.Ns, self.Ls)

    @property
    def β_shortrange(self) -> float:
        return self.α * self.rcut

    @property
    def β_longrange(self) -> float:
        return self.kmax / (2 * self.α)

    def background_vorticity_correction_factor(self) -> float:
        """⟨ψ⟩ = C·⟨ω⟩ with C = 1/(4α²) (``gaussian.jl:241``)."""
        return 1.0 / (4 * self.α * self.α)

    def __str__(self) -> str:
        return (
            f"GaussianSplitting(Ls={self.Ls}, α={self.α}, rcut={self.rcut}, Ns={self.Ns}, "
            f"kmax={self.kmax}, rtol={self.rtol})"
        )


def _as_periods(Ls) -> Tuple[float, float, float]:
    if np.isscalar(Ls):
        return (float(Ls),) * 3
    Ls = tuple(float(L) for L in Ls)
    if len(Ls) != 3:
        raise ValueError("Ls must have three entries")
    return Ls


def _as_dims(Ns) -> Tuple[int, int, int]:
    if np.isscalar(Ns):
        return (int(Ns),) * 3
    Ns = tuple(int(n) for n in Ns)
    if len(Ns) != 3:
        raise ValueError("Ns must have three entries")
    return Ns


# --------------------------------------------------------------------------- #
# Backends


def max_cutoff_distance_celllists(M: int, L: float) -> float:
    """rcut ≤ M/(2M+1)·L (``CellLists.jl:115``)."""
    return (M / (2 * M + 1)) * L


@dataclass(frozen=True)
class HIPShortRangeBackend:
    """Short-range pair sum on the GPU over a device-built cell list.

    Replaces ``CellListsBackend(ROCBackend(), nsubdiv)``; ``nsubdiv`` has the
    same meaning (cells of size ≈ rcut/nsubdiv, ``CellLists.jl:131-137``).
    """

    nsubdiv: int = 2
    device: int = 0

    def max_cutoff_distance(self, Ls) -> float:
        return max_cutoff_distance_celllists(self.nsubdiv, min(Ls))


@dataclass(frozen=True)
class HIPLongRangeBackend:
    """Long-range part on the GPU: hand-written NUFFT (ES kernel) and transforms.

    Replaces ``NonuniformFFTsBackend(ROCBackend(); σ, m)``. ``σ`` is the
    oversampling factor and ``m`` the kernel half-support, as in the reference
    (``nonuniformffts.jl:107-120``); the defaults give ~6-digit accuracy.
    ``fft`` selects the transforms: ``"fused"`` (hand-written FFTs fused with the
    pencil reorderings and the Fourier-space kernel; grids up to 6144 (FP32) or
    3072 (FP64) along x and y and half that along z, larger ones use rocFFT; on
    x-grids of 192, 384 or 768 cells (768: FP32 only) the inverse x transform also
    runs inside the interpolation), ``"separate"`` (the same
    transforms, interpolation in a kernel of its own) or ``"rocfft"`` (rocFFT 1D
    transforms with separate reordering and Fourier-space kernels).
    """

    σ: float = 1.5
    m: int = 4
    device: int = 0
    fft: str = "fused"

    def __post_init__(self):
        if self.fft not in ("fused", "separate", "rocfft"):
            raise ValueError(f"fft must be 'fused', 'separate' or 'rocfft', got {self.fft!r}")

    def has_real_to_complex(self) -> bool:
        return True


# --------------------------------------------------------------------------- #
# ParamsBiotSavart


_FLOAT_TYPES = {np.float64: "f64", np.float32: "f32", float: "f64"}


@dataclass(frozen=True)
class ParamsBiotSavart:
    """Parameters of the Ewald-split Biot–Savart evaluation (``params.jl:326-376``).

    ``T`` is the precision of the bulk arithmetic: ``np.float64`` runs the plain
    FP64 path, ``np.float32`` the FP32-bulk path (FP64 kept for node
    differencing, the local LIA log term and the local self terms). Inputs and
    outputs are float64 in both cases.
    """

    T: type
    Γ: float
    a: float
    Δ: float
    splitting: GaussianSplitting
    quadrature: object
    quadrature_near_singularity: object
    backend_short: HIPShortRangeBackend
    backend_long: HIPLongRangeBackend
    longrange_truncate_spherical: bool
    lia_segment_fraction: Optional[float]
    avoid_explicit_erf: bool
    rcut_sq: float = field(init=False)

    def __post_init__(self):
        rcut = self.splitting.rcut
        rcut_max = self.backend_short.max_cutoff_distance(self.Ls)
        if rcut > rcut_max:  # params.jl:100-104
            raise ValueError(
                f"cutoff distance `rcut = {rcut}` is larger than that allowed by the "
                f"{type(self.backend_short).__name__} backend (rcut_max = {rcut_max})"
            )
        if not (2 * rcut <= min(self.Ls)):  # params.jl:105-106
            raise ValueError(
                f"cutoff distance `rcut = {rcut}` is too large. It must be less than half "
                f"the cell unit size `L` in each direction: Ls = {self.Ls}."
            )
        if self.lia_segment_fraction is not None:
            g = float(self.lia_segment_fraction)
            if not (0 < g <= 1):
                raise ValueError("lia_segment_fraction must be in (0, 1]")
        object.__setattr__(self, "rcut_sq", rcut * rcut)  # params.jl:107

    # Convenience accessors (params.jl:412-419, 440-447)
    @property
    def Ls(self):
        return self.splitting.Ls

    @property
    def Ns(self):
        return self.splitting.Ns

    @property
    def α(self):
        return self.splitting.α

    @property
    def rcut(self):
        return self.splitting.rcut

    @property
    def precision(self) -> str:
        return _FLOAT_TYPES[self.T]

    def circulation(self) -> float:
        return self.Γ

    def periods(self):
        return self.Ls

    def maximum_wavenumber(self) -> float:
        return maximum_wavenumber(self.Ns, self.Ls)

    def __str__(self) -> str:
        return (
            f"ParamsBiotSavart{{{self.precision}}}(Γ={self.Γ}, a={self.a}, Δ={self.Δ}, "
            f"{self.splitting}, quad={self.quadrature}, backend_short={self.backend_short}, "
            f"backend_long={self.backend_long}, avoid_explicit_erf={self.avoid_explicit_erf})"
        )


def _make_params(
    T=np.float64,
    *,
    Γ,
    a,
    splitting: Optional[GaussianSplitting] = None,
    α=None,
    Ls=None,
    Ns=None,
    rcut=None,
    Δ: float = 0.25,
    quadrature=None,
    quadrature_near_singularity=None,
    backend_short: Optional[HIPShortRangeBackend] = None,
    backend_long: Optional[HIPLongRangeBackend] = None,
    longrange_truncate_spherical: bool = False,
    lia_segment_fraction: Optional[float] = None,
    avoid_explicit_erf: bool = True,
) -> ParamsBiotSavart:
    if T not in _FLOAT_TYPES:
        raise TypeError("T must be np.float64 or np.float32")
    if splitting is None:
        # Deprecated-but-supported keywords (params.jl:364-373, 432-437)
        if α is None or Ls is None or Ns is None or rcut is None:
            raise ValueError("pass `splitting`, or all of `α`, `Ls`, `Ns`, `rcut`")
        splitting = GaussianSplitting.create(Ls=Ls, α=α, rcut=rcut, Ns=Ns)
    if quadrature is None:
        quadrature = GaussLegendre(3)  # params.jl:340
    if quadrature_near_singularity is None:
        quadrature_near_singularity = quadrature
    if backend_short is None:
        backend_short = HIPShortRangeBackend(2)  # params.jl:392
    if backend_long is None:
        backend_long = HIPLongRangeBackend(1.5, 4)  # params.jl:395
    return ParamsBiotSavart(
        T=T,
        Γ=float(Γ),
        a=float(a),
        Δ=float(Δ),
        splitting=splitting,
        quadrature=quadrature,
        quadrature_near_singularity=quadrature_near_singularity,
        backend_short=backend_short,
        backend_long=backend_long,
        longrange_truncate_spherical=bool(longrange_truncate_spherical),
        lia_segment_fraction=None if lia_segment_fraction is None else float(lia_segment_fraction),
        avoid_explicit_erf=bool(avoid_explicit_erf),
    )


# Keyword constructor mirroring ``ParamsBiotSavart([T]; Γ, a, splitting, kws...)``
# (params.jl:337-376): ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=...).
ParamsBiotSavart.build = staticmethod(_make_params)
