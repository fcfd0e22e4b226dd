"""Learned correction of a cheap Ewald splitting (no counterpart in the reference).

The velocity of a cheap Gaussian splitting (β = 2 on a grid coarser than the
accurate one: short cut-off, few modes) is corrected by a small network that
predicts, per node, the difference to a converged splitting (β = 4 on the same
grid, accurate to ~6e-9; the reference's error model, ``gaussian.jl:45-62``,
puts β = 2 at ~1e-3) from local geometry:

* features (``vpbs_learned_features``, computed on the GPU): the node's local
  frame (t, n, b) from s′ and s″, its curvature, and the 16 nearest strands
  within twice the cheap cut-off, each as the foot of the perpendicular on the
  strand's tangent line plus the strand's tangent, in the node's frame, lengths
  in units of 1/α;
* target (``vpbs_learned_difference``): the exact converged − cheap velocity
  (FP64 pair sums over the converged cut-off, the kept Fourier field with the
  difference of the two Ewald filters), in the node's frame, in units of
  Γα/4π, so that the network is independent of Γ and of the length scale;
* network: 128 inputs → 3 × 128 SiLU → 3 outputs, trained with the library's
  own HIP forward, backward and Adam kernels (``vpbs_mlp_train``), run with
  BF16 MFMA inference (``vpbs_mlp_infer``);
* judge: every evaluation (``vpbs_compute_learned``) checks the corrected
  velocity against the converged splitting on a 1 % sample of nodes and falls
  back to the converged splitting on every node when the error exceeds the
  tolerance.

Training samples live in device buffers (``_lib.DeviceArray``, plain hipMalloc);
every computation on them runs in libvpbs.
"""

from __future__ import annotations

import json
import math
import pathlib
import time
from dataclasses import asdict, dataclass
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .filaments import FilamentBatch
from .params import GaussianSplitting, HIPLongRangeBackend, ParamsBiotSavart, maximum_wavenumber

__all__ = [
    "N_IN",
    "N_USED",
    "LearnedSettings",
    "CorrectionNetwork",
    "training_pairs",
    "TrainingSet",
    "SampleStats",
    "fit",
    "LearnedBiotSavart",
    "DEFAULT_WEIGHTS",
]

N_IN = 128   # feature row of the library
N_USED = 102  # 6 + 6·16 filled entries
HIDDEN, LAYERS, N_OUT = 128, 3, 3
DEFAULT_WEIGHTS = pathlib.Path(__file__).resolve().parent / "data" / "correction.npz"


@dataclass(frozen=True)
class LearnedSettings:
    """Cheap and converged splittings on the same grid, and the feature radius.

    ``Ns``: the cheap (coarse) Fourier grid; α = kmax/(2β), rcut = β/α for both
    β (``gaussian.jl:155-162``). ``beta_short``, when given, sets the cheap cut-off
    rcut = beta_short/α independently of the α of ``beta_cheap`` (the long-range
    truncation stays that of ``beta_cheap``, the short-range truncation becomes
    ~erfc(beta_short)). ``feature_factor``: feature radius in units of the cheap
    cut-off.
    """

    Ls: Tuple[float, float, float]
    Ns: Tuple[int, int, int]
    beta_cheap: float = 2.0
    beta_conv: float = 4.0
    feature_factor: float = 2.0
    beta_short: Optional[float] = None
    nufft_m: int = 4
    nufft_sigma: float = 1.5

    @staticmethod
    def create(Ls, Ns, beta_cheap=2.0, beta_conv=4.0, feature_factor=2.0, beta_short=None, nufft_m=4,
               nufft_sigma=1.5) -> "LearnedSettings":
        Ls = tuple(float(x) for x in np.broadcast_to(np.asarray(Ls, dtype=np.float64), (3,)))
        Ns = tuple(int(x) for x in np.broadcast_to(np.asarray(Ns), (3,)))
        bs = None if beta_short is None else float(beta_short)
        return LearnedSettings(Ls, Ns, float(beta_cheap), float(beta_conv), float(feature_factor), bs, int(nufft_m),
                               float(nufft_sigma))

    def to_meta(self) -> dict:
        return asdict(self)

    @staticmethod
    def from_meta(d: dict) -> "LearnedSettings":
        return LearnedSettings.create(**d)

    @property
    def cheap(self) -> GaussianSplitting:
        g = GaussianSplitting.create(Ls=self.Ls, β=self.beta_cheap, Ns=self.Ns)
        if self.beta_short is None:
            return g
        return GaussianSplitting.create(Ls=self.Ls, α=g.α, rcut=self.beta_short / g.α, Ns=self.Ns)

    @property
    def converged(self) -> GaussianSplitting:
        return GaussianSplitting.create(Ls=self.Ls, β=self.beta_conv, Ns=self.Ns)

    @property
    def feature_radius(self) -> float:
        return self.feature_factor * self.cheap.rcut

    def params(self, T=np.float32, Γ=1.0, a=1e-6, Δ=0.25, **kws) -> ParamsBiotSavart:
        """Parameters of the cheap context. Local segments are excluded from the pair sum and
        their long-range part removed (avoid_explicit_erf = false), in both precisions."""
        kws.setdefault("backend_long", HIPLongRangeBackend(σ=self.nufft_sigma, m=self.nufft_m))
        return ParamsBiotSavart.build(T, Γ=Γ, a=a, Δ=Δ, splitting=self.cheap, avoid_explicit_erf=False, **kws)


class CorrectionNetwork:
    """Parameters (flat FP32, ``include/vpbs.h`` layout), normalisation, output bound and metadata.

    ``out_bound``: per output, the largest |target| of the training set times ``BOUND_MARGIN``;
    the inference clamps the prediction to it (None: no bound).
    """

    BOUND_MARGIN = 1.5

    def __init__(self, params, in_shift, in_scale, out_scale, meta: Optional[dict] = None, out_bound=None):
        self.params = np.asarray(params, dtype=np.float32)
        self.in_shift = np.asarray(in_shift, dtype=np.float32)
        self.in_scale = np.asarray(in_scale, dtype=np.float32)
        self.out_scale = np.asarray(out_scale, dtype=np.float32)
        self.out_bound = None if out_bound is None else np.asarray(out_bound, dtype=np.float32)
        self.meta = dict(meta or {})

    @staticmethod
    def num_params() -> int:
        n, i = 0, N_IN
        for _ in range(LAYERS):
            n += HIDDEN * i + HIDDEN
            i = HIDDEN
        return n + N_OUT * HIDDEN + N_OUT

    @staticmethod
    def initial(seed: int = 0) -> "CorrectionNetwork":
        """He-uniform weights for SiLU, zero biases, no weights on the unused inputs."""
        rng = np.random.default_rng(seed)
        parts, fan = [], N_USED
        i = N_IN
        for l in range(LAYERS + 1):
            o = HIDDEN if l < LAYERS else N_OUT
            lim = math.sqrt((6.0 if l < LAYERS else 3.0) / fan)
            W = rng.uniform(-lim, lim, size=(o, i))
            if l == 0:
                W[:, N_USED:] = 0.0
            parts += [W.ravel(), np.zeros(o)]
            i = fan = HIDDEN
        return CorrectionNetwork(np.concatenate(parts), np.zeros(N_IN), np.ones(N_IN), np.ones(N_OUT))

    def set_normalisation(self, stats: "SampleStats"):
        """Per-feature mean and 1/std (constant features: shift only), per-output rms."""
        mean, std = stats.x_mean, stats.x_std
        self.in_shift = mean.astype(np.float32)
        self.in_scale = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-30), 1.0).astype(np.float32)
        self.out_scale = stats.t_rms.astype(np.float32)
        self.out_bound = (self.BOUND_MARGIN * stats.t_max).astype(np.float32)

    def to_device(self, device: int = 0) -> "_lib.MLP":
        net = _lib.MLP(N_IN, HIDDEN, LAYERS, N_OUT, device=device)
        net.set_params(self.params, self.in_shift, self.in_scale, self.out_scale)
        net.set_output_bound(self.out_bound)
        return net

    def save(self, path):
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        extra = {} if self.out_bound is None else {"out_bound": self.out_bound}
        np.savez(path, params=self.params, in_shift=self.in_shift, in_scale=self.in_scale, out_scale=self.out_scale,
                 meta=np.frombuffer(json.dumps(self.meta).encode(), dtype=np.uint8), **extra)

    @staticmethod
    def load(path=DEFAULT_WEIGHTS) -> "CorrectionNetwork":
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(bytes(z["meta"]).decode())
            bound = z["out_bound"] if "out_bound" in z.files else None
            return CorrectionNetwork(z["params"], z["in_shift"], z["in_scale"], z["out_scale"], meta, bound)


def _context(settings: LearnedSettings, T, Γ=1.0, device=0, **kws):
    p = settings.params(T, Γ=Γ, **kws)
    ctx = _lib.Context(p, device=device)
    conv = settings.converged
    return p, ctx, conv


class SampleStats:
    """Running per-column sums of features and targets (FP64, on the host)."""

    def __init__(self):
        self.n = 0
        self.sx = np.zeros(N_IN)
        self.sxx = np.zeros(N_IN)
        self.stt = np.zeros(N_OUT)
        self.tmax = np.zeros(N_OUT)

    def add(self, X: np.ndarray, T: np.ndarray):
        X = X.astype(np.float64)
        self.n += X.shape[0]
        self.sx += X.sum(axis=0)
        self.sxx += (X * X).sum(axis=0)
        self.stt += (T.astype(np.float64) ** 2).sum(axis=0)
        self.bound(T)

    def bound(self, T: np.ndarray):
        """Largest |target| per output (every row of T)."""
        self.tmax = np.maximum(self.tmax, np.abs(T.astype(np.float64)).max(axis=0, initial=0.0))

    @property
    def t_max(self):
        return self.tmax

    @property
    def x_mean(self):
        return self.sx / self.n

    @property
    def x_std(self):
        m = self.x_mean
        return np.sqrt(np.maximum(self.sxx / self.n - m * m, 0.0) * self.n / max(self.n - 1, 1))

    @property
    def t_rms(self):
        return np.sqrt(self.stt / self.n)


class TrainingSet:
    """Features X (n, 128) and targets T (n, 3), float32 DeviceArrays, with their statistics."""

    def __init__(self, X, T, stats: SampleStats):
        self.X, self.T, self.stats = X, T, stats

    def __len__(self):
        return self.X.shape[0]

    @staticmethod
    def concatenate(sets: Sequence["TrainingSet"]) -> "TrainingSet":
        st = SampleStats()
        for s in sets:
            st.n += s.stats.n
            st.sx += s.stats.sx
            st.sxx += s.stats.sxx
            st.stt += s.stats.stt
            st.tmax = np.maximum(st.tmax, s.stats.tmax)
        return TrainingSet(_lib.DeviceArray.concatenate([s.X for s in sets]),
                           _lib.DeviceArray.concatenate([s.T for s in sets]), st)


def training_pairs(tangles: Iterable[Sequence], settings: LearnedSettings, Γ: float = 1.0, device: int = 0,
                   log=None) -> TrainingSet:
    """Features and targets of every node of every tangle (device arrays (n, 128), (n, 3)).

    Each tangle is evaluated once with the cheap splitting in FP64 (point data built on the
    device), then the converged − cheap difference is computed exactly on every node. The
    per-column statistics (for the normalisation) come from every 8th row, the output bound from
    every row.
    """
    p, ctx, conv = _context(settings, np.float64, Γ=Γ, device=device)
    ctx.learned_setup(None, conv.α, conv.rcut, settings.feature_radius, 0.0, 1.0)
    Xs, Ts = [], []
    stats = SampleStats()
    for k, fs in enumerate(tangles):
        t0 = time.perf_counter()
        batch = fs if isinstance(fs, FilamentBatch) else FilamentBatch(fs)
        ctx.set_filaments(batch, p)
        ctx.compute(_lib.VELOCITY | _lib.KEEP_UHAT, fetch=False)
        X = _lib.DeviceArray((batch.Np, N_IN), np.float32, device)
        T = _lib.DeviceArray((batch.Np, N_OUT), np.float32, device)
        ctx.learned_features(X, None)
        ctx.learned_difference(None, T)
        stats.add(X.rows(8), T.rows(8))
        stats.bound(T.numpy())
        Xs.append(X)
        Ts.append(T)
        if log:
            log(f"pairs: tangle {k}: {batch.Np} nodes ({time.perf_counter() - t0:.1f} s)")
    ctx.close()
    out = TrainingSet(_lib.DeviceArray.concatenate(Xs, device), _lib.DeviceArray.concatenate(Ts, device), stats)
    for a in Xs + Ts:
        a.free()
    return out


def fit(net: CorrectionNetwork, data: TrainingSet, steps: int = 20000, batch: int = 8192, lr: float = 2e-3,
        lr_final: float = 2e-5, chunks: int = 20, seed: int = 0, device: int = 0, log=None) -> CorrectionNetwork:
    """Adam on the GPU (``vpbs_mlp_train``) with a learning rate decaying geometrically from
    ``lr`` to ``lr_final`` over ``chunks`` equal parts. Normalisation from the set's statistics."""
    X, T = data.X, data.T
    net.set_normalisation(data.stats)
    dev = net.to_device(device)
    per = max(1, steps // chunks)
    hist = []
    for c in range(chunks):
        lr_c = lr * (lr_final / lr) ** (c / max(1, chunks - 1))
        losses = dev.train(X, T, steps=per, batch=batch, lr=lr_c, seed=seed + c)
        hist.append(float(np.mean(losses[-max(1, per // 10):])))
        if log:
            log(f"fit: chunk {c}: lr {lr_c:.2e}, batch loss {hist[-1]:.4e}")
    net.params, net.in_shift, net.in_scale, net.out_scale = dev.get_params()
    net.meta.update({"steps": per * chunks, "batch": batch, "lr": lr, "lr_final": lr_final, "loss_history": hist})
    dev.close()
    return net


class LearnedBiotSavart:
    """Velocity on nodes with the cheap splitting plus the learned correction (FP32 bulk by default).

    ``velocity_on_nodes(fs)`` returns ``(v (Np, 3), report)``; the report gives the sampled
    errors of the corrected and the uncorrected velocity against the converged splitting and
    whether the correction was rejected (then v is the converged splitting's velocity).
    """

    def __init__(self, settings: LearnedSettings, network: CorrectionNetwork, T=np.float32, Γ: float = 1.0,
                 sample_fraction: float = 0.01, tol: float = 1e-4, seed: int = 0, device: int = 0, **kws):
        self.settings = settings
        self.params, self.ctx, conv = _context(settings, T, Γ=Γ, device=device, **kws)
        self.net = network.to_device(device)
        self.ctx.learned_setup(self.net, conv.α, conv.rcut, settings.feature_radius, sample_fraction, tol, seed)

    def velocity_on_nodes(self, fs, fetch: bool = True):
        batch = fs if isinstance(fs, FilamentBatch) else FilamentBatch(fs)
        self.ctx.set_filaments(batch, self.params)
        v, rep = self.ctx.compute_learned(fetch=fetch)
        return (v.T if v is not None else None), rep

    def close(self):
        self.ctx.close()
        self.net.close()
