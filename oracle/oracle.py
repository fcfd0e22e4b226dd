"""FP64 CPU oracle for the Ewald-split Biot–Savart evaluation.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker. The product path
(``vortexpasta.jl_amd``) never imports it.

The heavy loops live in ``vpbs_oracle.c`` (built into ``liboracle.so`` by the
Makefile next to this file); this module strings them together exactly as the
reference's ``_compute_on_nodes!`` does (``src/BiotSavart/BiotSavart.jl:515-612``):

    v = LIA (replaces) + pair sum − self-interaction + long range
        [+ background-vorticity correction of ψ, BiotSavart.jl:255-286]

Long range (``longrange/longrange.jl:102-224,488-495``; callbacks
``cache_common.jl:64-76``):

    ω̂(k) = (Γ/V) Σ_j q_j e^{-ik·s_j}          (type-1 NUFFT or exact sum)
    v̂ = i k × ω̂ / k²,  ψ̂ = ω̂ / k²,  k = 0 → 0
    × exp(-k²/4α²)  (Ewald Gaussian, gaussian.jl:233-237)
    u(x_i) = Σ_k û(k) e^{ik·x_i}                 (type-2 NUFFT or exact sum)

with the retained mode set of ``_ensure_hermitian_symmetry!``
(``longrange.jl:556-583``). The NUFFT restates the published algorithm of
the (un-vendored) NonuniformFFTs.jl: spreading with an "exponential of
semicircle" (ES) kernel of half-support m on a σ-oversampled grid, FFT, and
division by the kernel's Fourier transform. The FFT is numpy/scipy's
pocketfft.
"""

from __future__ import annotations

import ctypes
import math
import os
import pathlib
from typing import Dict, Optional

import numpy as np
import scipy.fft

_HERE = pathlib.Path(__file__).resolve().parent
_LIB = None

FLAG_VEL = 1
FLAG_PSI = 2
FLAG_AVOID_ERF = 4

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_up = ctypes.POINTER(ctypes.c_uint64)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library not built: {path} (run __graft_entry__.build())")
        L = ctypes.CDLL(str(path))
        i64, i32, d, u = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_uint
        L.or_shortrange.argtypes = [i64, _dp, _ip, i64, _dp, _dp, i32, _dp, d, d, d, i32, i32, u, _dp, _dp, _lp, _up,
                                    i64, i64]
        L.or_direct_open.argtypes = [i64, _dp, _ip, i64, _dp, _dp, i32, d, u, _dp, _dp]
        L.or_self_interaction.argtypes = [i64, _dp, _ip, i64, _dp, _dp, i32, _dp, d, d, u, _dp, _dp]
        L.or_local_term.argtypes = [i64, _dp, _dp, _dp, _dp, d, d, d, u, _dp, _dp]
        L.or_exact_type1.argtypes = [i64, _dp, _dp, _dp, _ip, d, _dp]
        L.or_exact_type2.argtypes = [i64, _dp, _dp, _ip, _dp, _dp]
        L.or_spread.argtypes = [i64, _dp, _dp, _dp, _ip, i32, d, _dp]
        L.or_interp.argtypes = [i64, _dp, _dp, _ip, i32, d, _dp, _dp]
        L.or_cell_dims.argtypes = [d, _dp, i32, _ip]
        L.or_local_integrals.argtypes = [i64, _dp, i32, _dp, _dp, _dp, i32, d, u, _dp, _dp]
        L.or_pair_hash.argtypes = [i64]
        L.or_pair_hash.restype = ctypes.c_uint64
        _LIB = L
    return _LIB


def _p(a, kind=_dp):
    return a.ctypes.data_as(kind) if a is not None else None


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# --------------------------------------------------------------------------- #
# NUFFT kernel choices (shared definition with the HIP library, restated here)


def es_beta(m: int, sigma: float) -> float:
    """ES kernel shape β = γ·π·m·(2 - 1/σ), γ = 0.98 (Barnett et al. 2019, §4.4)."""
    return 0.98 * math.pi * m * (2.0 - 1.0 / sigma)


def oversampled_size(N: int, sigma: float, m: int = 4) -> int:
    """Smallest multiple of 32 of the form 2^a·3^b·5^c that is ≥ ⌈σN⌉ (and ≥ 4m)."""
    n = int(math.ceil(sigma * N - 1e-9))
    while True:
        if n % 32 == 0:
            r = n
            for p in (2, 3, 5):
                while r % p == 0:
                    r //= p
            if r == 1:
                break
        n += 1
    if n < 4 * m:
        n = ((4 * m + 31) // 32) * 32
    return n


def es_fourier(k: np.ndarray, G: int, m: int, beta: float, nquad: int = 128) -> np.ndarray:
    """ψ̂(k) = 2m ∫_0^1 φ(z) cos(2πkmz/G) dz by Gauss–Legendre, φ(z) = e^{β(√(1-z²)-1)}."""
    x, w = np.polynomial.legendre.leggauss(nquad)
    z = (x + 1) / 2
    w = w / 2
    phi = np.exp(beta * (np.sqrt(1 - z * z) - 1))
    k = np.asarray(k, dtype=np.float64)
    return 2 * m * np.sum((w * phi)[None, :] * np.cos(2 * np.pi * k[:, None] * m * z[None, :] / G), axis=1)


def mode_limits(Ns):
    """Retained modes after the Hermitian clean-up (longrange.jl:556-583)."""
    kxmax = Ns[0] // 2 - 1
    kym = Ns[1] // 2 - 1 if Ns[1] % 2 == 0 else (Ns[1] - 1) // 2
    kzm = Ns[2] // 2 - 1 if Ns[2] % 2 == 0 else (Ns[2] - 1) // 2
    return kxmax, kym, kzm


def _fourier_operator(kx, ky, kz, Ls, alpha, quantity: str, truncate_kmax2: Optional[float]):
    """Returns (op_real_scalar, kvec) for ψ̂ = ω̂/k² or v̂ = i k × ω̂/k², both × e^{-k²/4α²}."""
    KX = 2 * np.pi * kx[None, None, :] / Ls[0]
    KY = 2 * np.pi * ky[None, :, None] / Ls[1]
    KZ = 2 * np.pi * kz[:, None, None] / Ls[2]
    k2 = KX * KX + KY * KY + KZ * KZ
    with np.errstate(divide="ignore"):
        k2inv = np.where(k2 == 0, 0.0, 1.0 / np.where(k2 == 0, 1.0, k2))
    ewald = np.exp(-k2 / (4 * alpha * alpha))
    scal = k2inv * ewald
    if truncate_kmax2 is not None:
        scal = np.where(k2 > truncate_kmax2, 0.0, scal)
    return scal, (KX, KY, KZ)


def _apply_operator(w, scal, kvec, quantity):
    """w: (3, nz, ny, nx) complex vorticity coefficients on retained modes."""
    if quantity == "streamfunction":
        return w * scal[None]
    KX, KY, KZ = kvec
    # i k × ω (longrange.jl:128-142)
    cx = KY * w[2] - KZ * w[1]
    cy = KZ * w[0] - KX * w[2]
    cz = KX * w[1] - KY * w[0]
    return 1j * np.stack([cx, cy, cz]) * scal[None]


def longrange(pd, params, quantities, method: str = "nufft", workers: int = 8) -> Dict[str, np.ndarray]:
    """Long-range velocity/streamfunction on nodes (LR part of compute_on_nodes!)."""
    L = lib()
    Ls = np.array(params.Ls, dtype=np.float64)
    Ns = np.array(params.Ns, dtype=np.int32)
    Gamma = params.Γ
    V = float(np.prod(Ls))
    pref = Gamma / V
    alpha = params.α
    kxmax, kym, kzm = mode_limits(params.Ns)
    kx = np.arange(0, kxmax + 1)
    ky = np.arange(-kym, kym + 1)
    kz = np.arange(-kzm, kzm + 1)
    trunc = None
    if params.longrange_truncate_spherical:
        trunc = params.maximum_wavenumber() ** 2
    nodes = _c(pd.nodes)
    pts = _c(pd.points)
    chg = _c(pd.charges)
    Np, Nq = nodes.shape[1], pts.shape[1]
    out = {}
    if method == "exact":
        Nkx = Ns[0] // 2 + 1
        uhat = np.zeros((3, Ns[2], Ns[1], Nkx, 2))
        L.or_exact_type1(Nq, _p(pts), _p(chg), _p(Ls), _p(Ns, _ip), pref, _p(uhat))
        uhat = uhat[..., 0] + 1j * uhat[..., 1]
        iy = ky % Ns[1]
        iz = kz % Ns[2]
        w = uhat[:, iz][:, :, iy][:, :, :, kx]
        for q in quantities:
            scal, kvec = _fourier_operator(kx, ky, kz, Ls, alpha, q, trunc)
            uq = _apply_operator(w, scal, kvec, q)
            full = np.zeros((3, Ns[2], Ns[1], Nkx), dtype=np.complex128)
            full[np.ix_(range(3), iz, iy, kx)] = uq
            buf = np.stack([full.real, full.imag], axis=-1).copy()
            res = np.zeros((3, Np))
            L.or_exact_type2(Np, _p(nodes), _p(Ls), _p(Ns, _ip), _p(buf), _p(res))
            out[q] = res
        return out

    # NUFFT path
    m = params.backend_long.m
    sigma = params.backend_long.σ
    beta = es_beta(m, sigma)
    G = np.array([oversampled_size(int(n), sigma, m) for n in Ns], dtype=np.int32)
    grid = np.zeros((3, G[2], G[1], G[0]))
    L.or_spread(Nq, _p(pts), _p(chg), _p(Ls), _p(G, _ip), m, beta, _p(grid))
    spec = scipy.fft.rfftn(grid, axes=(1, 2, 3), workers=workers)
    del grid
    phx = es_fourier(kx, int(G[0]), m, beta)
    phy = es_fourier(ky, int(G[1]), m, beta)
    phz = es_fourier(kz, int(G[2]), m, beta)
    D = 1.0 / (phz[:, None, None] * phy[None, :, None] * phx[None, None, :])
    iy = ky % G[1]
    iz = kz % G[2]
    w = spec[:, iz][:, :, iy][:, :, :, kx] * (pref * D)[None]
    del spec
    for q in quantities:
        scal, kvec = _fourier_operator(kx, ky, kz, Ls, alpha, q, trunc)
        uq = _apply_operator(w, scal, kvec, q) * D[None]
        full = np.zeros((3, G[2], G[1], G[0] // 2 + 1), dtype=np.complex128)
        full[np.ix_(range(3), iz, iy, kx)] = uq
        g2 = scipy.fft.irfftn(full, s=(int(G[2]), int(G[1]), int(G[0])), axes=(1, 2, 3), norm="forward",
                              workers=workers)
        del full
        g2 = np.ascontiguousarray(g2)
        res = np.zeros((3, Np))
        L.or_interp(Np, _p(nodes), _p(Ls), _p(G, _ip), m, beta, _p(g2), _p(res))
        out[q] = res
    return out


def _wavenumbers(Ns, Ls):
    """Reference wavenumber vectors (kx: r2c half axis; ky, kz: FFT order), cache_common.jl."""
    kx = 2 * np.pi / Ls[0] * np.arange(Ns[0] // 2 + 1)
    ky = 2 * np.pi / Ls[1] * np.where(np.arange(Ns[1]) < (Ns[1] + 1) // 2, np.arange(Ns[1]), np.arange(Ns[1]) - Ns[1])
    kz = 2 * np.pi / Ls[2] * np.where(np.arange(Ns[2]) < (Ns[2] + 1) // 2, np.arange(Ns[2]), np.arange(Ns[2]) - Ns[2])
    return kx, ky, kz


def longrange_fourier(pd, params, quantity: str = "velocity", method: str = "nufft", workers: int = 8) -> np.ndarray:
    """Unsmoothed long-range field in Fourier space, layout (3, N2, N1, N0//2 + 1) complex.

    What ``get_longrange_field_fourier`` (cache.jl:77-83) holds after
    ``compute_on_nodes!``: ω̂ = (Γ/V) Σ q e^{-ik·s} (type-1 NUFFT, longrange.jl:488-495),
    then ψ̂ = ω̂/k² (longrange.jl:116-124) or v̂ = i k × ω̂/k² (longrange.jl:128-142),
    without the Ewald factor (applied only when interpolating, cache_common.jl:64-76).
    Modes outside the retained set (longrange.jl:556-583) are zero. ``quantity``:
    "vorticity", "streamfunction" or "velocity".
    """
    L = lib()
    Ls = np.array(params.Ls, dtype=np.float64)
    Ns = np.array(params.Ns, dtype=np.int32)
    pref = params.Γ / float(np.prod(Ls))
    kxmax, kym, kzm = mode_limits(params.Ns)
    kx = np.arange(0, kxmax + 1)
    ky = np.arange(-kym, kym + 1)
    kz = np.arange(-kzm, kzm + 1)
    pts = _c(pd.points)
    chg = _c(pd.charges)
    Nq = pts.shape[1]
    Nkx = Ns[0] // 2 + 1
    iy = ky % Ns[1]
    iz = kz % Ns[2]
    if method == "exact":
        uh = np.zeros((3, Ns[2], Ns[1], Nkx, 2))
        L.or_exact_type1(Nq, _p(pts), _p(chg), _p(Ls), _p(Ns, _ip), pref, _p(uh))
        uh = uh[..., 0] + 1j * uh[..., 1]
        w = uh[:, iz][:, :, iy][:, :, :, kx]
    else:
        m = params.backend_long.m
        sigma = params.backend_long.σ
        beta = es_beta(m, sigma)
        G = np.array([oversampled_size(int(n), sigma, m) for n in Ns], dtype=np.int32)
        grid = np.zeros((3, G[2], G[1], G[0]))
        L.or_spread(Nq, _p(pts), _p(chg), _p(Ls), _p(G, _ip), m, beta, _p(grid))
        spec = scipy.fft.rfftn(grid, axes=(1, 2, 3), workers=workers)
        del grid
        D = 1.0 / (es_fourier(kz, int(G[2]), m, beta)[:, None, None] * es_fourier(ky, int(G[1]), m, beta)[None, :, None]
                   * es_fourier(kx, int(G[0]), m, beta)[None, None, :])
        w = spec[:, kz % G[2]][:, :, ky % G[1]][:, :, :, kx] * (pref * D)[None]
        del spec
    if quantity == "vorticity":
        u = w
    else:
        KX = 2 * np.pi * kx[None, None, :] / Ls[0]
        KY = 2 * np.pi * ky[None, :, None] / Ls[1]
        KZ = 2 * np.pi * kz[:, None, None] / Ls[2]
        k2 = KX * KX + KY * KY + KZ * KZ
        scal = np.where(k2 == 0, 0.0, 1.0 / np.where(k2 == 0, 1.0, k2))
        if params.longrange_truncate_spherical:
            scal = np.where(k2 > params.maximum_wavenumber() ** 2, 0.0, scal)
        u = _apply_operator(w, scal, (KX, KY, KZ), quantity)
    full = np.zeros((3, Ns[2], Ns[1], Nkx), dtype=np.complex128)
    full[np.ix_(range(3), iz, iy, kx)] = u
    return full


def energy_spectrum(uhat: np.ndarray, Ls, nk: int, dk: float, quantity: str = "velocity") -> np.ndarray:
    """Energy spectrum from a half-spectrum field (Diagnostics/spectra.jl:89-196).

    Bin n = trunc(|k|/Δk + 1/2) (0-based), E[n] += f·|û|²/Δk with f = 1/2 on the
    kx = 0 plane and 1 elsewhere (Hermitian symmetry); from vorticity |ω̂|²/k².
    """
    Ns = (2 * (uhat.shape[3] - 1), uhat.shape[2], uhat.shape[1])
    kx, ky, kz = _wavenumbers(Ns, np.asarray(Ls, dtype=np.float64))
    k2 = kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + kz[:, None, None] ** 2
    u2 = np.sum(np.abs(uhat) ** 2, axis=0)
    if quantity == "vorticity":
        u2 = u2 * np.where(k2 == 0, 1.0, 1.0 / np.where(k2 == 0, 1.0, k2))
    elif quantity != "velocity":
        raise ValueError("the energy spectrum needs the velocity or the vorticity")
    factor = np.where(kx == 0, 0.5, 1.0)[None, None, :]
    n = np.trunc(np.sqrt(k2) / dk + 0.5).astype(np.int64)
    E = np.zeros(nk)
    sel = n < nk
    np.add.at(E, n[sel], (factor * u2 / dk)[sel])
    return E


def band_field(uhat: np.ndarray, params, nodes: np.ndarray, kmin: float, kmax: float,
               method: str = "nufft", workers: int = 8) -> np.ndarray:
    """Σ over kmin² ≤ k² ≤ kmax² of û(k) e^{ik·x} at the nodes (the full Hermitian sum).

    The field behind FourierBandForcingBS (fourier_band_bs.jl; band selection as
    SyntheticFields.set_wavevector_band!, fourier.jl:241-266) and, with kmin = 0,
    the low-pass velocity of SmallScaleDissipationBS
    (dissipation_bs_small_scales.jl:109-128). ``method``: "exact" (direct sum) or
    "nufft" (deconvolution, inverse FFT, ES interpolation: the library's algorithm).
    """
    L = lib()
    Ls = np.array(params.Ls, dtype=np.float64)
    Ns = np.array(params.Ns, dtype=np.int32)
    kx, ky, kz = _wavenumbers(Ns, Ls)
    k2 = kx[None, None, :] ** 2 + ky[None, :, None] ** 2 + kz[:, None, None] ** 2
    mask = (k2 >= kmin * kmin) & (k2 <= kmax * kmax)
    u = uhat * mask[None]
    nodes = _c(nodes)
    Np = nodes.shape[1]
    res = np.zeros((3, Np))
    if method == "exact":
        buf = np.stack([u.real, u.imag], axis=-1).copy()
        L.or_exact_type2(Np, _p(nodes), _p(Ls), _p(Ns, _ip), _p(buf), _p(res))
        return res
    m = params.backend_long.m
    sigma = params.backend_long.σ
    beta = es_beta(m, sigma)
    G = np.array([oversampled_size(int(n), sigma, m) for n in Ns], dtype=np.int32)
    kxmax, kym, kzm = mode_limits(params.Ns)
    qx = np.arange(0, kxmax + 1)
    qy = np.arange(-kym, kym + 1)
    qz = np.arange(-kzm, kzm + 1)
    D = 1.0 / (es_fourier(qz, int(G[2]), m, beta)[:, None, None] * es_fourier(qy, int(G[1]), m, beta)[None, :, None]
               * es_fourier(qx, int(G[0]), m, beta)[None, None, :])
    full = np.zeros((3, G[2], G[1], G[0] // 2 + 1), dtype=np.complex128)
    full[np.ix_(range(3), qz % G[2], qy % G[1], qx)] = u[:, qz % Ns[2]][:, :, qy % Ns[1]][:, :, :, qx] * D[None]
    g2 = scipy.fft.irfftn(full, s=(int(G[2]), int(G[1]), int(G[0])), axes=(1, 2, 3), norm="forward", workers=workers)
    del full
    g2 = np.ascontiguousarray(g2)
    L.or_interp(Np, _p(nodes), _p(Ls), _p(G, _ip), m, beta, _p(g2), _p(res))
    return res


def local_cutoffs(params, integration_cutoff=None):
    if integration_cutoff is None:
        rpsi = math.exp(params.Δ - 1) * params.a / 2  # local_integrals.jl:118-119
        return rpsi * math.exp(1.0), rpsi
    return float(integration_cutoff), float(integration_cutoff)


def compute(pd, params, quantities=("velocity",), LIA=True, shortrange=True, longrange_on=True,
            lr_method="nufft", use_celllist=True, integration_cutoff=None, neighbour_stats=False,
            workers: int = 8):
    """Oracle of compute_on_nodes! on linear point data. Returns dict of (3, Np) arrays.

    ``LIA`` ∈ {True, False, "only"} as in ``BiotSavart.jl:515-612``.
    """
    L = lib()
    Ls = np.array(params.Ls, dtype=np.float64)
    nodes = _c(pd.nodes)
    pts = _c(pd.points)
    chg = _c(pd.charges)
    idx = _c(pd.node_idx_prev, np.int32)
    Np, Nq = nodes.shape[1], pts.shape[1]
    flags = 0
    if "velocity" in quantities:
        flags |= FLAG_VEL
    if "streamfunction" in quantities:
        flags |= FLAG_PSI
    if params.avoid_explicit_erf:
        flags |= FLAG_AVOID_ERF
    v = np.zeros((3, Np))
    psi = np.zeros((3, Np))
    stats = None
    if shortrange:
        if LIA is True or LIA == "only":
            cv, cpsi = local_cutoffs(params, integration_cutoff)
            L.or_local_term(Np, _p(_c(pd.sp)), _p(_c(pd.spp)), _p(_c(pd.dlm)), _p(_c(pd.dlp)), params.Γ, cv, cpsi,
                            flags, _p(v), _p(psi))
        if LIA != "only":
            counts = np.zeros(Np, dtype=np.int64) if neighbour_stats else None
            hashes = np.zeros(Np, dtype=np.uint64) if neighbour_stats else None
            L.or_shortrange(Np, _p(nodes), _p(idx, _ip), Nq, _p(pts), _p(chg), pd.nquad, _p(Ls), params.α,
                            params.rcut, params.Γ, params.backend_short.nsubdiv, int(use_celllist), flags,
                            _p(v), _p(psi), _p(counts, _lp), _p(hashes, _up), 0, -1)
            L.or_self_interaction(Np, _p(nodes), _p(idx, _ip), Nq, _p(pts), _p(chg), pd.nquad, _p(Ls), params.α,
                                  params.Γ, flags, _p(v), _p(psi))
            if neighbour_stats:
                stats = (counts, hashes)
    if longrange_on and LIA != "only":
        lr = longrange(pd, params, quantities, method=lr_method, workers=workers)
        if "velocity" in lr:
            v += lr["velocity"]
        if "streamfunction" in lr:
            psi += lr["streamfunction"]
    if shortrange and LIA != "only":
        local_integrals(pd, params, flags, v, psi, nodes, periodic=True)
        if "streamfunction" in quantities:
            psi += background_vorticity_psi(pd, params)[:, None]
    out = {}
    if "velocity" in quantities:
        out["velocity"] = v
    if "streamfunction" in quantities:
        out["streamfunction"] = psi
    if neighbour_stats:
        out["stats"] = stats
    return out


def _subsegment_points(pd, params, fold: bool):
    """Quadrature points/charges on the non-LIA parts of the two segments of each node.

    Segment i-1 over ζ ∈ [0, 1-γ] and segment i over ζ ∈ [γ, 1]
    (local_integrals.jl:103-104,224-241), node-major layout.
    """
    g = params.lia_segment_fraction
    b = pd.batch
    zs, ws = params.quadrature_near_singularity.rule()
    nq = len(zs)
    N = b.Ns[b.fid]
    pts, chg = [], []
    for seg, (lo, hi) in (((b.local - 1) % N, (0.0, 1.0 - g)), (b.local, (g, 1.0))):
        dt = (b.knot(b.fid, seg + 1) - b.knot(b.fid, seg)) * (hi - lo)
        for z, w in zip(zs, ws):
            y = np.full(b.Np, lo + (hi - lo) * z)
            pts.append(b.evaluate(b.fid, seg, y, 0))
            chg.append((w * dt)[:, None] * b.evaluate(b.fid, seg, y, 1))
    P = np.stack(pts, axis=1).reshape(-1, 3).T  # (3, 2nq*Np), node-major
    C = np.stack(chg, axis=1).reshape(-1, 3).T
    if fold:
        for d in range(3):
            L = params.Ls[d]
            P[d] = np.mod(P[d], L)
            P[d][P[d] >= L] -= L
    return 2 * nq, np.ascontiguousarray(P), np.ascontiguousarray(C)


def local_integrals(pd, params, flags, v, psi, nodes, periodic=True):
    if params.lia_segment_fraction is None:
        return
    nsub, P, C = _subsegment_points(pd, params, fold=periodic)
    Ls = np.array(params.Ls, dtype=np.float64)
    lib().or_local_integrals(nodes.shape[1], _p(nodes), nsub, _p(P), _p(C), _p(Ls), int(periodic), params.Γ, flags,
                             _p(v), _p(psi))


def background_vorticity_psi(pd, params) -> np.ndarray:
    """ψ_back = -(Γ/V) Σ_f offset_f / (4α²) when filaments are unclosed (BiotSavart.jl:255-286)."""
    Ls = np.array(params.Ls)
    ptot = pd.batch.Xoff.sum(axis=0)
    if np.allclose(ptot + Ls, Ls, rtol=np.sqrt(np.finfo(float).eps), atol=0):
        return np.zeros(3)
    wmean = ptot * (params.Γ / np.prod(Ls))
    return -wmean * params.splitting.background_vorticity_correction_factor()


def compute_open_domain(pd, params, quantities=("velocity",), integration_cutoff=None):
    """NoSplitting reference: LIA + direct sum over non-local segments, no periodicity."""
    L = lib()
    nodes = _c(pd.batch.X.T)  # unfolded nodes
    # unfolded points: fold undone by evaluating again is not needed; use batch positions
    b = pd.batch
    nq = pd.nquad
    zs, _ = params.quadrature.rule()
    seg = np.repeat(b.local, nq)
    fq = np.repeat(b.fid, nq)
    pts = _c(b.evaluate(fq, seg, np.tile(zs, b.Np), 0).T)
    chg = _c(pd.charges)
    idx = _c(pd.node_idx_prev, np.int32)
    Np, Nq = nodes.shape[1], pts.shape[1]
    flags = (FLAG_VEL if "velocity" in quantities else 0) | (FLAG_PSI if "streamfunction" in quantities else 0)
    v = np.zeros((3, Np))
    psi = np.zeros((3, Np))
    cv, cpsi = local_cutoffs(params, integration_cutoff)
    L.or_local_term(Np, _p(_c(pd.sp)), _p(_c(pd.spp)), _p(_c(pd.dlm)), _p(_c(pd.dlp)), params.Γ, cv, cpsi, flags,
                    _p(v), _p(psi))
    L.or_direct_open(Np, _p(nodes), _p(idx, _ip), Nq, _p(pts), _p(chg), nq, params.Γ, flags, _p(v), _p(psi))
    local_integrals(pd, params, flags, v, psi, nodes, periodic=False)
    out = {}
    if flags & FLAG_VEL:
        out["velocity"] = v
    if flags & FLAG_PSI:
        out["streamfunction"] = psi
    return out


def neighbour_pairs_bruteforce(pd, params):
    """Per-node (count, hash) of accepted pairs by brute force (CellLists test, cell_lists.jl:43-60)."""
    L = lib()
    Ls = np.array(params.Ls, dtype=np.float64)
    nodes, pts, chg = _c(pd.nodes), _c(pd.points), _c(pd.charges)
    idx = _c(pd.node_idx_prev, np.int32)
    Np, Nq = nodes.shape[1], pts.shape[1]
    counts = np.zeros(Np, dtype=np.int64)
    hashes = np.zeros(Np, dtype=np.uint64)
    flags = FLAG_VEL | (FLAG_AVOID_ERF if params.avoid_explicit_erf else 0)
    v = np.zeros((3, Np))
    L.or_shortrange(Np, _p(nodes), _p(idx, _ip), Nq, _p(pts), _p(chg), pd.nquad, _p(Ls), params.α, params.rcut,
                    params.Γ, params.backend_short.nsubdiv, 0, flags, _p(v), None, _p(counts, _lp), _p(hashes, _up),
                    0, -1)
    return counts, hashes


def energy_from_streamfunction(pd, psi, params, open_domain=False):
    """E = Γ/(2V)·Σ_i ψ_i·s′_i·(t_{i+1} - t_{i-1})/2 (Diagnostics/energy.jl:86-119).

    Open domains (test/ring.jl:118) use Γ/2 instead of Γ/(2V).
    """
    b = pd.batch
    dtp = b.segment_dt()
    dtm = dtp[pd.node_idx_prev]
    w = (dtp + dtm) / 2
    integ = np.sum(np.sum(psi * pd.sp, axis=0) * w)
    if open_domain:
        return params.Γ / 2 * integ
    return params.Γ / (2 * np.prod(params.Ls)) * integ
