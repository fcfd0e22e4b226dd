"""Public Biot–Savart API: init_cache, compute_on_nodes_, velocity_on_nodes(_).

Mirrors ``src/BiotSavart/BiotSavart.jl``:

* ``init_cache(params)`` — ``cache.jl:259-268``; builds the GPU context
  (cell-list geometry, NUFFT plans) once and reuses it across calls.
* ``compute_on_nodes_(fields, cache, fs; LIA, shortrange, longrange,
  integration_cutoff)`` — ``BiotSavart.jl:395-404,515-612``. ``fields`` is a
  dict with keys ``"velocity"`` and/or ``"streamfunction"``, each a list with
  one ``(N_i, 3)`` float64 array per filament, overwritten in place (a
  trailing underscore stands for Julia's ``!``). ``LIA`` is ``True``,
  ``False`` or ``"only"`` (``Val(true)``, ``Val(false)``, ``Val(:only)``).
* ``velocity_on_nodes_(vs, cache, fs)`` / ``velocity_on_nodes(cache, fs)`` —
  ``BiotSavart.jl:94-161``.

Work split: the spline coefficients of the filaments are solved on the host
(as the reference's ``Filaments`` module does); the point data (quadrature
points, charges, derivatives on nodes, ``add_point_charges!``,
``pointdata_impl.jl``) is built from them on the GPU by default
(``pointdata="device"``, ``vpbs_set_filaments``) or on the host and uploaded
(``pointdata="host"``, ``vpbs_set_points``, the reference's split); the
short-range and long-range parts run on the GPU through libvpbs; the
``lia_segment_fraction`` local integrals run on the GPU with device point data
and on the host otherwise; the background-vorticity correction of ψ for
unclosed filaments stays on the host (``BiotSavart.jl:255-286,570-580``).
Errors from the library raise ``VpbsError``; invalid arguments raise
``ValueError`` like the reference's ``ArgumentError``/``DimensionMismatch``.
"""

from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from .filaments import ClosedFilament, FilamentBatch
from .local_integrals import add_local_integrals
from .params import ParamsBiotSavart
from .pointdata import PointData, add_point_charges

__all__ = [
    "BiotSavartCache",
    "get_longrange_field_fourier",
    "init_cache",
    "compute_on_nodes_",
    "velocity_on_nodes_",
    "velocity_on_nodes",
    "background_vorticity_correction",
]

_FIELDS = ("velocity", "streamfunction")


class BiotSavartCache:
    """Parameters + GPU context + last point data (``cache.jl:238-250``).

    ``pointdata``: ``"device"`` builds the point data on the GPU from the
    filament splines, ``"host"`` builds it with :func:`add_point_charges` and
    uploads it (then kept in ``self.pointdata``).
    ``keep_fourier``: keep the unsmoothed long-range field in Fourier space
    after each evaluation, as the reference's long-range cache does (read by
    :func:`get_longrange_field_fourier`, the energy spectrum and the forcing
    and dissipation terms).
    """

    def __init__(self, params: ParamsBiotSavart, device: int = 0, pointdata: str = "device",
                 keep_fourier: bool = True):
        if pointdata not in ("device", "host"):
            raise ValueError("pointdata must be 'device' or 'host'")
        self.params = params
        self.ctx = _lib.Context(params, device=device)
        self.pointdata_mode = pointdata
        self.keep_fourier = keep_fourier
        self.pointdata: Optional[PointData] = None
        self.batch: Optional[FilamentBatch] = None
        self.timings: Dict[str, float] = {}

    def close(self):
        self.ctx.close()


def init_cache(params: ParamsBiotSavart, device: int = 0, pointdata: str = "device",
               keep_fourier: bool = True) -> BiotSavartCache:
    return BiotSavartCache(params, device=device, pointdata=pointdata, keep_fourier=keep_fourier)


def get_longrange_field_fourier(cache: BiotSavartCache) -> dict:
    """Long-range field in Fourier space of the last evaluation (``cache.jl:77-83``).

    Returns ``{"field": (ux, uy, uz), "wavenumbers": (kx, ky, kz), "state": quantity}``
    with each component of shape ``(N2, N1, N0//2+1)`` (indexed [kz, ky, kx]) and
    ``quantity`` "velocity" or "streamfunction" (unsmoothed, as the reference's
    ``LongRangeCacheState`` with ``smoothing_scale = 0``).
    """
    q, u = cache.ctx.get_uhat()
    if q is None:
        raise RuntimeError("no long-range Fourier field: evaluate with keep_fourier=True and the long range on")
    Ls = cache.params.Ls
    Ns = cache.params.Ns
    kx = 2 * np.pi / Ls[0] * np.arange(Ns[0] // 2 + 1)
    ky = 2 * np.pi / Ls[1] * np.fft.fftfreq(Ns[1], 1.0 / Ns[1])
    kz = 2 * np.pi / Ls[2] * np.fft.fftfreq(Ns[2], 1.0 / Ns[2])
    return {"field": (u[0], u[1], u[2]), "wavenumbers": (kx, ky, kz), "state": q}


def background_vorticity_correction(batch: FilamentBatch, params: ParamsBiotSavart) -> np.ndarray:
    """Uniform ψ shift −(Γ/V)Σ offsets/(4α²) when the filaments' mean vorticity is non-zero."""
    Ls = np.asarray(params.Ls, dtype=np.float64)
    ptot = batch.Xoff.sum(axis=0)
    if np.allclose(ptot + Ls, Ls, rtol=np.sqrt(np.finfo(float).eps), atol=0):
        return np.zeros(3)
    wmean = ptot * (params.Γ / np.prod(Ls))
    return -wmean * params.splitting.background_vorticity_correction_factor()


def _flags(fields, LIA, shortrange, longrange) -> int:
    f = 0
    if "velocity" in fields:
        f |= _lib.VELOCITY
    if "streamfunction" in fields:
        f |= _lib.STREAMFUNCTION
    if LIA is False:
        f |= _lib.LIA_OFF
    elif LIA == "only":
        f |= _lib.LIA_ONLY
    elif LIA is not True:
        raise ValueError("LIA must be True, False or 'only'")
    if not shortrange:
        f |= _lib.NO_SHORTRANGE
    if not longrange:
        f |= _lib.NO_LONGRANGE
    return f


def compute_on_nodes_(
    fields: Dict[str, List[np.ndarray]],
    cache: BiotSavartCache,
    fs: Sequence[ClosedFilament],
    *,
    LIA=True,
    shortrange: bool = True,
    longrange: bool = True,
    integration_cutoff: Optional[float] = None,
    neighbour_stats: bool = False,
):
    """Compute velocity and/or streamfunction on filament nodes (``BiotSavart.jl:395-404``).

    ``fs`` is a list of filaments or a :class:`FilamentBatch` built from them
    (whose spline coefficients are then reused, as the reference's filaments
    carry theirs between evaluations).
    """
    if isinstance(fs, FilamentBatch):
        batch, fs = fs, fs.fs
    else:
        batch = None
    if not fields:
        raise ValueError("fields must contain 'velocity' and/or 'streamfunction'")
    unknown = set(fields) - set(_FIELDS)
    if unknown:
        raise ValueError(f"some of these fields were not recognised: {sorted(unknown)}")
    for name, us in fields.items():  # BiotSavart.jl:171-173
        if len(us) != len(fs) or any(np.shape(u) != (len(f), 3) for u, f in zip(us, fs)):
            raise ValueError(f"wrong dimensions of output vector '{name}'")
    params = cache.params
    if len(fs) == 0:
        # no filaments: nothing to evaluate (the output lists are empty as well)
        return fields
    if batch is None:
        batch = FilamentBatch(fs)
    if cache.pointdata_mode == "device":
        cache.pointdata = None
        cache.ctx.set_filaments(batch, params)
    else:
        pd = add_point_charges(fs, params, batch)
        cache.pointdata = pd
        cache.ctx.set_points(pd)
    flags = _flags(fields, LIA, shortrange, longrange)
    if neighbour_stats:
        flags |= _lib.NEIGHBOUR_STATS
    if cache.keep_fourier and longrange:
        flags |= _lib.KEEP_UHAT
    cache.batch = batch
    v, psi = cache.ctx.compute(flags, integration_cutoff)
    cache.timings = cache.ctx.timings()
    if shortrange and LIA != "only":
        # BiotSavart.jl:570-580; with device point data the local integrals were
        # added on the device by vpbs_compute
        if cache.pointdata_mode == "host":
            add_local_integrals(v, psi, batch, params)
        if psi is not None:
            psi += background_vorticity_correction(batch, params)[:, None]
    for name, arr in (("velocity", v), ("streamfunction", psi)):
        if name in fields:
            lin = arr.T
            for u, a, n in zip(fields[name], batch.offs, batch.Ns):
                u[...] = lin[a : a + n]
    return fields


def velocity_on_nodes_(vs: List[np.ndarray], cache: BiotSavartCache, fs, **kws):
    compute_on_nodes_({"velocity": vs}, cache, fs, **kws)
    return vs


def velocity_on_nodes(cache: BiotSavartCache, fs, **kws) -> List[np.ndarray]:
    lens = fs.Ns if isinstance(fs, FilamentBatch) else [len(f) for f in fs]
    vs = [np.empty((int(n), 3)) for n in lens]
    return velocity_on_nodes_(vs, cache, fs, **kws)
