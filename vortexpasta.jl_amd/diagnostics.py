"""Diagnostics consuming the Biot–Savart outputs.

* Energy spectrum of the long-range field, mirroring ``Diagnostics.init_spectrum`` /
  ``energy_spectrum`` / ``energy_spectrum!`` (``src/Diagnostics/spectra.jl:49-115``):
  computed on the GPU from the Fourier field the last evaluation kept
  (``vpbs_energy_spectrum``), without copying that field to the host.
* Kinetic energy from streamfunction values on nodes,
  ``kinetic_energy_from_streamfunction`` without a quadrature rule
  (``src/Diagnostics/energy.jl:104-119``): ``E = Γ/(2V) Σ_i ψ_i·s′_i (t_{i+1} − t_{i−1})/2``.
"""

from __future__ import annotations

from typing import Tuple

import numpy as np

from .biotsavart import BiotSavartCache

__all__ = ["init_spectrum", "energy_spectrum", "energy_spectrum_", "kinetic_energy_from_streamfunction"]


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


def kinetic_energy_from_streamfunction(batch, psi: np.ndarray, Γ: float, Ls=None, sp: np.ndarray = None) -> float:
    """Kinetic energy per unit mass from streamfunction values on nodes (``energy.jl:104-119``).

    ``batch``: the filaments (a ``FilamentBatch``); ``psi``: (3, Np) or (Np, 3)
    streamfunction on the nodes in batch order; ``Ls``: periods (``None`` or an
    infinite period gives V = 1, ``energy.jl:70-75``); ``sp``: s′ on the nodes,
    (3, Np), evaluated from the splines when not given. The trapezoidal weight of
    node i is (t_{i+1} − t_{i−1})/2 with the filament's knots.
    """
    psi = np.asarray(psi, dtype=np.float64)
    if psi.shape == (batch.Np, 3):
        psi = psi.T
    if psi.shape != (3, batch.Np):
        raise ValueError(f"psi must have shape (3, {batch.Np}), got {psi.shape}")
    if sp is None:
        sp = batch.evaluate(batch.fid, batch.local, np.zeros(batch.Np), deriv=1).T
    dt = (batch.knot(batch.fid, batch.local + 1) - batch.knot(batch.fid, batch.local - 1)) / 2
    V = 1.0
    if Ls is not None and all(np.isfinite(L) for L in Ls):
        V = float(np.prod(Ls))
    return Γ / (2 * V) * float(np.sum(np.einsum("dn,dn->n", psi, sp) * dt))
