"""Configuration 3 (3000 random rings, ≈1.0M nodes, 512³ Fourier grid, β = 3.5) on the GPU
against one full evaluation of the FP64 CPU oracle on the same point data.

Bars: plain FP64 path ≤ 1e-12 relative L2 for velocity and streamfunction and
per-node neighbour sets (accepted-pair counts and hashes) identical to the
oracle's cell lists; FP32-bulk path within 1e-6 relative L2 for velocity and
streamfunction and 1e-6 relative for the energy (the ES NUFFT at m = 4, σ = 1.5
is itself accurate to ~4e-7; FP32 rounding adds ~1e-7).

The oracle run takes ~1 min on 16 host cores and ~36 GB of host memory.
"""
import math
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

Q = ("velocity", "streamfunction")


def _workers():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def cfg3(vp, O):
    fs = vp.tangles.config_tangle(3)
    spl = vp.GaussianSplitting.create(Ls=2 * math.pi, β=3.5, Ns=512)
    p64 = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, Δ=0.25, splitting=spl)
    p32 = vp.ParamsBiotSavart.build(np.float32, Γ=1.0, a=1e-6, Δ=0.25, splitting=spl)
    batch = vp.FilamentBatch(fs)
    pd = vp.add_point_charges(fs, p64, batch=batch)
    print(f"\n  config 3: {len(fs)} rings, Np = {pd.Np}, Nq = {pd.Nq}; running the oracle ...", flush=True)
    want = O.compute(pd, p64, Q, neighbour_stats=True, workers=_workers())
    print("  oracle done", flush=True)
    return pd, p32, p64, want


def _run(vp, params, pd, flags):
    ctx = vp._lib.Context(params)
    ctx.set_points(pd)
    v, psi = ctx.compute(flags, fetch=True)
    stats = ctx.neighbour_stats() if flags & vp._lib.NEIGHBOUR_STATS else None
    ctx.close()
    return v, psi, stats


def test_cfg3_fp64_vs_oracle(vp, cfg3):
    pd, _, p64, want = cfg3
    assert 950_000 < pd.Np < 1_060_000
    L = vp._lib
    v, psi, (counts, hashes) = _run(vp, p64, pd, L.VELOCITY | L.STREAMFUNCTION | L.NEIGHBOUR_STATS)
    ev, ep = rel_l2(v, want["velocity"]), rel_l2(psi, want["streamfunction"])
    print(f"  FP64 vs oracle: velocity {ev:.3e}, streamfunction {ep:.3e}", flush=True)
    assert ev <= 1e-12 and ep <= 1e-12
    oc, oh = want["stats"]
    assert np.array_equal(counts, oc) and np.array_equal(hashes, oh)
    print(f"  neighbour sets identical: {int(counts.sum())} accepted pairs", flush=True)


def test_cfg3_fp32_vs_oracle(vp, cfg3):
    pd, p32, p64, want = cfg3
    L = vp._lib
    v, psi, _ = _run(vp, p32, pd, L.VELOCITY | L.STREAMFUNCTION)
    ev, ep = rel_l2(v, want["velocity"]), rel_l2(psi, want["streamfunction"])
    E = vp.diagnostics.kinetic_energy_from_streamfunction(pd.batch, psi, p64.Γ, p64.Ls, sp=pd.sp)
    E0 = vp.diagnostics.kinetic_energy_from_streamfunction(pd.batch, want["streamfunction"], p64.Γ, p64.Ls, sp=pd.sp)
    eE = abs(E - E0) / abs(E0)
    print(f"  FP32 vs oracle: velocity {ev:.3e}, streamfunction {ep:.3e}, energy {eE:.3e}", flush=True)
    assert ev < 1e-6 and ep < 1e-6 and eE < 1e-6
