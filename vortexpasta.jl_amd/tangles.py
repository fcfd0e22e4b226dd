"""Synthetic vortex tangles used by tests and benchmarks.

Random rings follow the reference benchmark generator
(``benchmark/benchmarks.jl:48-66``): radius R uniform in [R_min, R_max],
centre uniform in the box, uniformly random orientation (a random unit
quaternion), N = ⌈2πR/Δs⌉ nodes at τ_i = (2i-1)/(2N) of the circle. Periodic
lines follow ``PredefinedCurves/periodic_line.jl`` with Kelvin-wave
perturbations. The random streams are numpy's (seeded), not the reference's
StableRNG, so the tangles are the same *distribution*, not the same samples.
"""

from __future__ import annotations

import math
from typing import List

import numpy as np

from .filaments import ClosedFilament, PeriodicLine, Ring, define_curve, init

__all__ = ["random_rotation", "random_rings", "single_ring", "kelvin_lines", "config_tangle"]

TWO_PI = 2 * math.pi


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniformly distributed rotation matrix (normalised Gaussian quaternion)."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
        ]
    )


def single_ring(R: float = math.pi / 3, N: int = 64, L: float = TWO_PI) -> List[ClosedFilament]:
    """One ring of radius R centred in [0, L)³ (``test/ring.jl:13-28`` with translate = L/2)."""
    S = define_curve(Ring(), scale=R, translate=L / 2)
    return [init(S, N)]


def random_rings(nrings: int, R_min: float, R_max: float, ds: float, L=TWO_PI, seed: int = 42) -> List[ClosedFilament]:
    rng = np.random.default_rng(seed)
    Ls = np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))
    fs = []
    for _ in range(nrings):
        R = R_min + rng.random() * (R_max - R_min)
        N = max(5, math.ceil(TWO_PI * R / ds))
        centre = Ls * rng.random(3)
        rot = random_rotation(rng)
        S = define_curve(Ring(), scale=R, translate=centre, rotate=rot)
        fs.append(init(S, N))
    return fs


def kelvin_lines(nlines: int, ds: float, L: float = TWO_PI, modes=range(1, 9), amplitude_total: float = 0.05 * TWO_PI,
                 seed: int = 7) -> List[ClosedFilament]:
    """Periodic lines along x, y and z with Kelvin-wave perturbations.

    Mode k has amplitude ∝ 1/k, the amplitudes summing to ``amplitude_total``;
    phases are random. Orientation cycles x, y, z, and alternates sign so that
    the total end-to-end offset (mean vorticity) is small but the lines are
    individually unclosed.
    """
    rng = np.random.default_rng(seed)
    ks = np.array(list(modes), dtype=np.float64)
    amps = (1 / ks) / np.sum(1 / ks) * amplitude_total / L  # in units of the line period (scaled by L)
    N = max(5, math.ceil(L / ds))
    perms = [np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]]),  # line along x
             np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]]),  # along y
             np.eye(3)]  # along z
    fs = []
    for n in range(nlines):
        phx = rng.random(len(ks)) * TWO_PI
        phy = rng.random(len(ks)) * TWO_PI

        def xf(t, phx=phx):
            return np.sum(amps[:, None] * np.cos(TWO_PI * ks[:, None] * np.atleast_1d(t)[None, :] + phx[:, None]), axis=0)

        def yf(t, phy=phy):
            return np.sum(amps[:, None] * np.sin(TWO_PI * ks[:, None] * np.atleast_1d(t)[None, :] + phy[:, None]), axis=0)

        axis = n % 3
        sign = 1 if (n // 3) % 2 == 0 else -1
        rot = perms[axis] @ np.diag([1, sign, sign]).astype(float)
        centre = L * rng.random(3)
        S = define_curve(PeriodicLine(x=xf, y=yf), scale=L, translate=centre, rotate=rot)
        fs.append(init(S, N))
    return fs


def config_tangle(cfg: int, seed: int = 42) -> List[ClosedFilament]:
    """Filaments of the benchmark configurations (1: one ring; 2: 64 rings ≈16k nodes;
    3: ≈3000 rings ≈1M nodes; 4: ≈256k nodes; 5: 8192 rings with radii U[0.2, 0.8] plus 256
    Kelvin-wave lines along x, y and z, node spacing 2π/1024, ≈4.4M nodes)."""
    if cfg == 1:
        return single_ring(math.pi / 3, 64)
    if cfg == 2:
        return random_rings(64, 0.5, 1.5, TWO_PI / 256, seed=seed)
    if cfg == 3:
        return random_rings(3000, 0.3, 1.0, TWO_PI / 512, seed=seed)
    if cfg == 4:
        return random_rings(1000, 0.5, 1.5, TWO_PI / 256, seed=seed)
    if cfg == 5:
        return random_rings(8192, 0.2, 0.8, TWO_PI / 1024, seed=seed) + kelvin_lines(256, TWO_PI / 1024, seed=seed + 1)
    raise ValueError(f"unknown configuration {cfg}")
