"""Energy spectrum of the long-range field (device-side ``Diagnostics`` consumer).

Mirrors ``Diagnostics.init_spectrum`` / ``energy_spectrum`` / ``energy_spectrum!``
(``src/Diagnostics/spectra.jl:49-115``): the spectrum is computed on the GPU from
the Fourier field the last evaluation kept (``vpbs_energy_spectrum``), without
copying that field to the host.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np

from .biotsavart import BiotSavartCache

__all__ = ["init_spectrum", "energy_spectrum", "energy_spectrum_"]


def init_spectrum(cache: BiotSavartCache) -> Tuple[np.ndarray, np.ndarray]:
    """Wavenumbers ``0:Δk:kx_max`` of the r2c axis and an uninitialised spectrum (``spectra.jl:49-80``)."""
    Ls, Ns = cache.params.Ls, cache.params.Ns
    dk = 2 * np.pi / Ls[0]
    ks = dk * np.arange(Ns[0] // 2 + 1)
    return ks, np.empty_like(ks)


def energy_spectrum_(Ek: np.ndarray, ks: np.ndarray, cache: BiotSavartCache):
    """Fill ``Ek`` (same length as ``ks``; ``ks[0] == 0``, constant step) on the device (``spectra.jl:89-123``)."""
    if len(Ek) != len(ks):
        raise ValueError("incompatible dimensions of vectors")
    if ks[0] != 0:
        raise ValueError("output wavenumbers should include k = 0")
    state = cache.ctx.uhat_quantity()
    if state != "velocity":  # spectra.jl:92-107 (vorticity state is not kept by this library)
        raise RuntimeError(f"the current state of the long-range cache ({state}) is currently not supported")
    dk = float(ks[1] - ks[0])
    Ek[...] = cache.ctx.energy_spectrum(len(ks), dk)
    return ks, Ek


def energy_spectrum(cache: BiotSavartCache) -> Tuple[np.ndarray, np.ndarray]:
    ks, Ek = init_spectrum(cache)
    return energy_spectrum_(Ek, ks, cache)
