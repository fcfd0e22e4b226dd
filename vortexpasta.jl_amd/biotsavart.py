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
``lia_segment_fraction`` local integrals and the background-vorticity correction
of ψ for unclosed filaments stay on the host (``BiotSavart.jl:255-286,570-580``).
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
    """

    def __init__(self, params: ParamsBiotSavart, device: int = 0, pointdata: str = "device"):
        if pointdata not in ("device", "host"):
            raise ValueError("pointdata must be 'device' or 'host'")
        self.params = params
        self.ctx = _lib.Context(params, device=device)
        self.pointdata_mode = pointdata
        self.pointdata: Optional[PointData] = None
        self.timings: Dict[str, float] = {}

    def close(self):
        self.ctx.close()


def init_cache(params: ParamsBiotSavart, device: int = 0, pointdata: str = "device") -> BiotSavartCache:
    return BiotSavartCache(params, device=device, pointdata=pointdata)


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
    """Compute velocity and/or streamfunction on filament nodes (``BiotSavart.jl:395-404``)."""
    if not fields:
        raise ValueError("fields must contain 'velocity' and/or 'streamfunction'")
    unknown = set(fields) - set(_FIELDS)
    if unknown:
        raise ValueError(f"some of these fields were not recognised: {sorted(unknown)}")
    for name, us in fields.items():  # BiotSavart.jl:171-173
        if len(us) != len(fs) or any(np.shape(u) != (len(f), 3) for u, f in zip(us, fs)):
            raise ValueError(f"wrong dimensions of output vector '{name}'")
    params = cache.params
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
    v, psi = cache.ctx.compute(flags, integration_cutoff)
    cache.timings = cache.ctx.timings()
    if shortrange and LIA != "only":
        # host-side terms, as in the reference (BiotSavart.jl:570-580)
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
    vs = [np.empty((len(f), 3)) for f in fs]
    return velocity_on_nodes_(vs, cache, fs, **kws)
