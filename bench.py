#!/usr/bin/env python3
"""Benchmark: Ewald-split Biot–Savart velocity on ~1M filament nodes (one MI355X).

Workload (default): 3000 random vortex rings, radii U[0.3, 1.0], node spacing
2π/512 (≈1.0M nodes, ≈3.0M quadrature points), periodic box [0, 2π)³, Fourier
grid 512³ (σ = 1.5 → 768³ oversampled NUFFT grid), β = 3.5 (α = 36.43,
rcut = 0.0961), quintic splines, GL(3), Γ = 1, a = 1e-6, Δ = 1/4.

A step is one velocity evaluation (velocity_on_nodes!) of the whole tangle with
the point data already resident in HBM: cell-list build, short-range pair sum,
LIA and self terms, NUFFT spreading, FFTs, Fourier-space kernel, interpolation
and the final sum. ``value`` is node-velocities/s of the FP32-bulk path; the
plain FP64 path is timed too, and the FP32 result is compared with the FP64
one (the FP64 path matches the CPU oracle to ≤1e-12 in the parity tests).

Prints ONE JSON line. ``--gpus`` above 1 is not supported (the library runs on one GPU).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

import __graft_entry__ as ge

METRIC = "Biot–Savart node-velocities/s at 1M nodes; rel L2 error vs FP64 oracle"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
VALU_PEAK_TFS = 78.6  # MI355X vector FP64, and FP32 without packed instructions (compiled out of all device code)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class DeviceWorkload:
    """The filaments only (point data built on the device by vpbs_set_filaments)."""

    def __init__(self, batch):
        self.batch = batch
        self.Np = batch.Np
        self.Nq = 0


def build_workload(vp, cfg: int, sigma: float = 1.5, m: int = 4, fft: str = "fused", host_points: bool = True):
    fs = vp.tangles.config_tangle(cfg)
    Ns = {1: 64, 2: 128, 3: 512, 4: 256, 5: 1024}[cfg]
    spl = vp.GaussianSplitting.create(Ls=2 * math.pi, β=3.5, Ns=Ns)
    bl = vp.HIPLongRangeBackend(σ=sigma, m=m, fft=fft)
    p32 = vp.ParamsBiotSavart.build(np.float32, Γ=1.0, a=1e-6, Δ=0.25, splitting=spl, backend_long=bl)
    p64 = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, Δ=0.25, splitting=spl, backend_long=bl)
    batch = vp.FilamentBatch(fs)
    if not host_points:
        return fs, spl, p32, p64, DeviceWorkload(batch)
    t = time.perf_counter()
    pd = vp.add_point_charges(fs, p64, batch=batch)
    pd.host_build_s = time.perf_counter() - t
    return fs, spl, p32, p64, pd


def time_path(vp, params, pd, steps: int, warmup: int, extra_flags: int = 0):
    ctx = vp._lib.Context(params)
    ctx.set_points(pd)  # H2D, outside the timed region
    flags = vp._lib.VELOCITY | extra_flags
    for _ in range(warmup):
        ctx.compute(flags, fetch=False)
    acc = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.compute(flags, fetch=False)  # returns after both streams are synchronised
        t = ctx.timings()
        for k, v in t.items():
            if isinstance(v, float):
                acc[k] = acc.get(k, 0.0) + v
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    stages = {k: v / steps for k, v in acc.items() if k != "total"}
    v, _ = ctx.compute(flags, fetch=True)
    grid, cells = ctx.sizes()
    ctx.close()
    return ms, stages, v, grid, cells


def gpu_fields(vp, params, pd):
    """Velocity and streamfunction of one (untimed) evaluation in the precision of ``params``."""
    ctx = vp._lib.Context(params)
    ctx.set_points(pd)
    out = ctx.compute(vp._lib.VELOCITY | vp._lib.STREAMFUNCTION, fetch=True)
    ctx.close()
    return out


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def errors_vs_oracle(vp, fields, ref, pd, p64):
    """Relative L2 errors of velocity, streamfunction and energy against the oracle's fields.

    The energy is ``kinetic_energy_from_streamfunction`` on the nodes
    (``src/Diagnostics/energy.jl:104-119``) of each streamfunction.
    """
    v, psi = fields
    out = {"velocity": rel_l2(v, ref["velocity"])}
    if psi is not None and "streamfunction" in ref:
        E = vp.diagnostics.kinetic_energy_from_streamfunction(pd.batch, psi, p64.Γ, p64.Ls, sp=pd.sp)
        E0 = vp.diagnostics.kinetic_energy_from_streamfunction(pd.batch, ref["streamfunction"], p64.Γ, p64.Ls,
                                                              sp=pd.sp)
        out.update({"streamfunction": rel_l2(psi, ref["streamfunction"]), "energy": abs(E - E0) / abs(E0),
                    "energy_oracle": E0})
    return out


def copy_calibration(nbytes: int = 1 << 30, reps: int = 5):
    """Plain copy kernel of known size, run alone so that rocprofv3 --pmc can scale its counters.

    Launches an elementwise kernel reading ``nbytes`` and writing ``nbytes``
    (``torch.neg`` on float32) ``reps`` times.
    """
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"read_bytes": nbytes, "write_bytes": nbytes, "launches": reps}}),
          flush=True)


PMC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                           "pmc_traffic_rings3000_1M_nodes_fp32.txt")


def measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 --pmc summary (None if absent).

    The summary lists one row per kernel, ``| kernel | dispatches | FETCH_SIZE KiB |
    WRITE_SIZE KiB | HBM bytes |``, with the copy-kernel calibration already applied.
    """
    try:
        with open(PMC_SUMMARY) as f:
            for line in f:
                cols = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cols) >= 5 and cols[0] == kernel:
                    return float(cols[4])
    except OSError:
        return None
    return None


def time_point_data(vp, params, pd, reps: int = 3):
    """Device point-data build (vpbs_set_filaments) against the host add_point_charges.

    Reports the kernel time (HIP events), the wall time of the call (upload of
    nodes, knots and spline coefficients included) and whether the device point
    data is bit-identical to the host-built one.
    """
    ctx = vp._lib.Context(params)
    ctx.set_filaments(pd.batch, params)  # first call allocates
    walls, kern = [], []
    for _ in range(reps):
        t = time.perf_counter()
        ctx.set_filaments(pd.batch, params)
        walls.append(time.perf_counter() - t)
        kern.append(ctx.timings()["point_data"])
    dev = ctx.get_points()
    ctx.close()
    same = all(np.array_equal(dev[k], getattr(pd, k))
               for k in ("nodes", "node_idx_prev", "sp", "spp", "dlm", "dlp", "points", "charges"))
    return {"device_kernel_ms": round(float(np.median(kern)), 4),
            "device_call_ms": round(float(np.median(walls)) * 1e3, 3),
            "host_add_point_charges_s": round(pd.host_build_s, 3),
            "bit_identical_to_host": bool(same)}


def time_end_to_end(vp, params, pd, reps: int = 3):
    """Wall time of one evaluation including the host side and PCIe, FP32 path.

    device: upload of nodes, knots and spline coefficients (vpbs_set_filaments),
    compute, copy of the velocities back; host: the reference's split, with
    add_point_charges on the host, upload of the point data, compute, copy back.
    The spline coefficients are solved once beforehand (filament state).
    """
    ctx = vp._lib.Context(params)
    res = {}
    for mode in ("device", "host"):
        walls = []
        for r in range(reps + 1):
            t = time.perf_counter()
            if mode == "device":
                ctx.set_filaments(pd.batch, params)
            else:
                ctx.set_points(vp.add_point_charges(pd.batch.fs, params, batch=pd.batch))
            v, _ = ctx.compute(vp._lib.VELOCITY, fetch=True)
            if r:
                walls.append(time.perf_counter() - t)
        res[f"{mode}_point_data_ms"] = round(float(np.median(walls)) * 1e3, 2)
    ctx.close()
    return res


def stage_bytes(stage: str, s: int, G, Ns, Np: int, Nq: int, slab: bool = False) -> float:
    """Algorithmic HBM bytes per launch of each long-range stage (velocity only).

    G: oversampled grid, Ns: Fourier grid, s: bytes per real of the compute type.
    spread 3Gs + 6Nq·s (grid written once, points read once); FFT forward
    3Gs + 6Zs (real grid in, z-pencils of the retained (kx, ky) out); Fourier
    kernel 12Zs (z-pencils in and out, 3 complex components); FFT backward
    6Zs + 3Gs; interpolation 3Gs + 6Np·s, with G = ∏G_d and Z = nkx·nky·G_2 the
    complex values per component of the z-pencils (nkx = N_0/2, nky = 2⌊(N_1-1)/2⌋+1).
    With the slab interpolation (``slab``) the inverse x transform moves from the
    backward stage into the interpolation: backward 6Zs + 6X·s (x2 in, x1 out),
    interpolation 6X·s + 6Np·s, X = nkx·G_1·G_2 retained x-modes per component.
    """
    Gc = float(np.prod(G))
    nkx = Ns[0] // 2
    nky = 2 * ((Ns[1] - 1) // 2) + 1
    Z = float(nkx) * nky * G[2]
    X = float(nkx) * G[1] * G[2]
    return {
        "spread": 3 * Gc * s + 6 * Nq * s,
        "fft_forward": 3 * Gc * s + 6 * Z * s,
        "fourier": 12 * Z * s,
        "fft_backward": 6 * Z * s + (6 * X * s if slab else 3 * Gc * s),
        "interp": (6 * X * s if slab else 3 * Gc * s) + 6 * Np * s,
    }[stage]


def pair_kernel_report(vp, params, pd, stage_ms: float, reps: int = 5):
    """Short-range pair kernel against the VALU peak.

    One untimed evaluation with VPBS_NEIGHBOUR_STATS gives the candidate sources
    tested (the cells of every node's (2M+1)³ neighbourhood that its rcut sphere
    reaches) and the accepted pairs;
    flops = 11 per candidate (FP64 separation, fold, r²) + 40 per accepted pair
    (sqrt, reciprocal, erfc, exp, weights, cross product, accumulation). The
    kernel is timed alone (short range only, HIP events) and beside the long range
    (``stage_ms``, from the timed evaluations).
    """
    L = vp._lib
    ctx = L.Context(params)
    ctx.set_points(pd)
    ctx.compute(L.VELOCITY | L.NEIGHBOUR_STATS, fetch=False)
    cand = int(ctx.timings()["pair_candidates"])
    counts, _ = ctx.neighbour_stats()
    acc = int(counts.sum())
    ctx.compute(L.VELOCITY | L.NO_LONGRANGE, fetch=False)
    alone = 0.0
    for _ in range(reps):
        ctx.compute(L.VELOCITY | L.NO_LONGRANGE, fetch=False)
        alone += ctx.timings()["pair"] / reps
    ctx.close()
    flops = 11.0 * cand + 40.0 * acc
    tf = flops / (alone * 1e-3) / 1e12
    return {"bound": "valu", "candidates": cand, "accepted": acc,
            "flops_model": "11 per candidate + 40 per accepted pair", "flops": flops,
            "ms_alone": round(alone, 4), "ms_beside_long_range": round(stage_ms, 4),
            "achieved_tflops": round(tf, 3), "peak_tflops": VALU_PEAK_TFS, "frac": round(tf / VALU_PEAK_TFS, 4)}


def oracle_run(O, pd, params):
    """One full FP64 oracle evaluation of the velocity (timed: the CPU baseline) and one of the
    streamfunction (untimed), on all host cores the process may use (at most 16)."""
    import resource

    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    t = time.perf_counter()
    ref = O.compute(pd, params, ("velocity",), workers=cores)
    t_v = time.perf_counter() - t
    log(f"oracle velocity: {t_v:.1f} s")
    t = time.perf_counter()
    ref["streamfunction"] = O.compute(pd, params, ("streamfunction",), workers=cores)["streamfunction"]
    log(f"oracle streamfunction: {time.perf_counter() - t:.1f} s")
    peak_gb = resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6
    cb = {"value": pd.Np / t_v, "unit": "node-velocities/s", "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
          "kind": "port",
          "sample": f"one full velocity evaluation of the same {pd.Np}-node workload by the FP64 C/numpy oracle "
                    f"(cell-list pair sum, LIA, self terms, ES-kernel NUFFT with scipy FFTs), {t_v:.1f} s",
          "seconds_per_eval": t_v, "peak_host_memory_GB": round(peak_gb, 1),
          "reference": "reference CPU path not timed: Julia unavailable"}
    return ref, cb


def cpu_baseline_sampled(O, pd, params, target_s: float = 12.0):
    """Time the FP64 CPU oracle on a bounded sample and scale to node-velocities/s (for
    configurations whose full oracle evaluation does not fit in the bench's time).

    Sample: short-range pair sum for a contiguous block of nodes (whole rings),
    spreading/interpolation for a contiguous block of points/nodes, one 3D r2c
    FFT of one velocity component on the full oversampled grid (scaled ×6 for
    the 3 forward + 3 backward transforms), LIA + self terms on all nodes.
    """
    import scipy.fft

    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    L = O.lib()
    Np, Nq = pd.Np, pd.Nq
    nodes = O._c(pd.nodes)
    pts = O._c(pd.points)
    chg = O._c(pd.charges)
    idx = O._c(pd.node_idx_prev, np.int32)
    Ls = np.array(params.Ls)
    v = np.zeros((3, Np))
    flags = O.FLAG_VEL | (O.FLAG_AVOID_ERF if params.avoid_explicit_erf else 0)

    def sr(i1):
        t = time.perf_counter()
        L.or_shortrange(Np, O._p(nodes), O._p(idx, O._ip), Nq, O._p(pts), O._p(chg), pd.nquad, O._p(Ls), params.α,
                        params.rcut, params.Γ, params.backend_short.nsubdiv, 1, flags, O._p(v), None, None, None, 0,
                        i1)
        return time.perf_counter() - t

    t_probe = sr(min(Np, 2000))  # includes the cell-list build
    n_sr = int(min(Np, max(2000, 2000 * (0.3 * target_s) / max(t_probe, 1e-6))))
    t_sr = sr(n_sr)
    t_sr_full = t_sr * Np / n_sr
    t = time.perf_counter()
    cv, cpsi = O.local_cutoffs(params)
    L.or_local_term(Np, O._p(O._c(pd.sp)), O._p(O._c(pd.spp)), O._p(O._c(pd.dlm)), O._p(O._c(pd.dlp)), params.Γ, cv,
                    cpsi, flags, O._p(v), None)
    L.or_self_interaction(Np, O._p(nodes), O._p(idx, O._ip), Nq, O._p(pts), O._p(chg), pd.nquad, O._p(Ls), params.α,
                          params.Γ, flags, O._p(v), None)
    t_local = time.perf_counter() - t
    # long range
    m, sigma = params.backend_long.m, params.backend_long.σ
    beta = O.es_beta(m, sigma)
    G = np.array([O.oversampled_size(int(n), sigma, m) for n in params.Ns], dtype=np.int32)
    grid = np.empty((3, G[2], G[1], G[0]))
    grid.fill(1.0)  # fault the pages in once (a CPU code reuses its grid across evaluations)
    t = time.perf_counter()
    grid.fill(0.0)  # per-evaluation zeroing of the spreading grid
    t_zero = time.perf_counter() - t
    nq_s = min(Nq, 300000)
    t = time.perf_counter()
    L.or_spread(nq_s, O._p(O._c(pts[:, :nq_s])), O._p(O._c(chg[:, :nq_s])), O._p(Ls), O._p(G, O._ip), m, beta,
                O._p(grid))
    t_spread = (time.perf_counter() - t) * Nq / nq_s + t_zero
    np_s = min(Np, 100000)
    out = np.zeros((3, np_s))
    t = time.perf_counter()
    L.or_interp(np_s, O._p(O._c(nodes[:, :np_s])), O._p(Ls), O._p(G, O._ip), m, beta, O._p(grid), O._p(out))
    t_interp = (time.perf_counter() - t) * Np / np_s
    t = time.perf_counter()
    spec = scipy.fft.rfftn(grid[0], workers=cores)
    t_fft = (time.perf_counter() - t) * 6
    del spec, grid
    total = t_sr_full + t_local + t_spread + t_fft + t_interp
    sample = (f"FP64 oracle, {cores} threads: pair sum on {n_sr} of {Np} nodes, spread {nq_s} of {Nq} points "
              f"(+ grid zero-fill), "
              f"interp {np_s} nodes, one {G[0]}^3 r2c FFT (x6), LIA+self on all nodes; scaled to one evaluation "
              f"of {total:.1f} s")
    return {"value": Np / total, "unit": "node-velocities/s", "cores": cores, "kind": "port", "sample": sample,
            "seconds_per_eval": total, "reference": "reference CPU path not timed: Julia unavailable",
            "breakdown_s": {"pairs": t_sr_full, "local_self": t_local, "spread": t_spread, "fft": t_fft,
                            "interp": t_interp}}


def learned_settings(vp, net, spl):
    """Cheap/converged settings of a trained network for this box: the betas, feature factor and
    NUFFT settings it was trained with, on half the workload's Fourier grid."""
    L = vp.learned
    meta = dict(net.meta.get("settings", {}))
    meta.update({"Ls": spl.Ls, "Ns": tuple(n // 2 for n in spl.Ns)})
    return L.LearnedSettings.from_meta(meta)


def time_learned(vp, pd, spl, steps: int, warmup: int, weights=None):
    """The cheap splitting (coarse grid, FP32 bulk) plus the learned correction with its 1 %
    sampled check (vpbs_compute_learned), timed per call (wall clock, the check and any fallback
    included), and the same cheap splitting without the network. Returns None when no trained
    network is available."""
    L = vp.learned
    path = pathlib.Path(weights) if weights else L.DEFAULT_WEIGHTS
    if not path.exists():
        log(f"learned correction: no weights at {path}")
        return None
    net = L.CorrectionNetwork.load(path)
    st = learned_settings(vp, net, spl)
    lb = L.LearnedBiotSavart(st, net, T=np.float32, sample_fraction=0.01, tol=1e-4)
    lb.ctx.set_filaments(pd.batch, lb.params)  # point data of the same filaments, built on the device
    for _ in range(warmup):
        lb.ctx.compute_learned(fetch=False)
    walls, reps, stages = [], [], {}
    for _ in range(steps):
        t = time.perf_counter()
        _, rep = lb.ctx.compute_learned(fetch=False)
        walls.append(time.perf_counter() - t)
        reps.append(rep)
        for k, v in lb.ctx.timings().items():
            if isinstance(v, float) and k != "total":
                stages[k] = stages.get(k, 0.0) + v / steps
    v, rep = lb.ctx.compute_learned(fetch=True)
    stats = lb.ctx.learned_stats()
    lb.close()
    # the cheap splitting alone (same context settings, no network, no check)
    ms_c, st_c, v_c, _, _ = time_path(vp, st.params(np.float32), pd, steps, warmup)
    ms = float(np.mean(walls)) * 1e3
    cheap = st.cheap
    return {"ms_per_step": ms, "value": pd.Np / (ms * 1e-3), "velocity": v, "velocity_cheap": v_c,
            "cheap": {"Ns": list(st.Ns), "beta_long": cheap.β_longrange, "beta_short": cheap.β_shortrange,
                      "alpha": cheap.α, "rcut": cheap.rcut, "nufft_m": st.nufft_m, "nufft_sigma": st.nufft_sigma,
                      "ms_per_step_without_network": ms_c,
                      "stages_ms_without_network": {k: round(v, 4) for k, v in st_c.items()}},
            "converged": {"beta": st.beta_conv, "alpha": st.converged.α, "rcut": st.converged.rcut},
            "sampled_check": {"fraction": 0.01, "tol": 1e-4,
                              "err_corrected_mean": float(np.mean([r["err_corrected"] for r in reps])),
                              "err_uncorrected_mean": float(np.mean([r["err_uncorrected"] for r in reps])),
                              "fallbacks": int(sum(r["fallback"] for r in reps)), "calls": stats["calls"]},
            "stages_ms": {k: round(v, 4) for k, v in stages.items()},
            "network": {"file": os.path.relpath(path, ge.ROOT),
                        "held_out": net.meta.get("held_out", {})}}


def time_advance(vp, pd, spl, steps: int, warmup: int, dt: float, weights=None, T=np.float32):
    """SSPRK33 steps with the filaments on the device (vpbs_advance, 3 velocity evaluations per
    step): with a trained network every evaluation is the learned one (FP32 bulk); with
    ``weights`` "none" the plain path in precision T."""
    L = vp.learned
    path = pathlib.Path(weights) if weights and weights != "none" else L.DEFAULT_WEIGHTS
    net = L.CorrectionNetwork.load(path) if weights != "none" and path.exists() else None
    if net is not None:
        st = learned_settings(vp, net, spl)
        lb = L.LearnedBiotSavart(st, net, T=np.float32, sample_fraction=0.01, tol=1e-4)
        ctx, params = lb.ctx, lb.params
    else:
        params = vp.ParamsBiotSavart.build(T, Γ=1.0, a=1e-6, Δ=0.25, splitting=spl)
        ctx, lb = vp._lib.Context(params), None
    ctx.set_filaments(pd.batch, params)
    if warmup:
        ctx.advance(dt, warmup)
    t0 = time.perf_counter()
    ctx.advance(dt, steps)
    wall = time.perf_counter() - t0
    out = {"steps": steps, "evaluations": 3 * steps, "dt": dt, "seconds": wall,
           "ms_per_evaluation": wall * 1e3 / (3 * steps), "learned": net is not None,
           "precision": "f32" if net is not None or T == np.float32 else "f64"}
    if lb is not None:
        out["learned_stats"] = ctx.learned_stats()
        lb.close()
    else:
        ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=[1, 2, 3, 4, 5])
    ap.add_argument("--sigma", type=float, default=1.5, help="NUFFT oversampling factor")
    ap.add_argument("--m", type=int, default=4, help="NUFFT kernel half-support")
    ap.add_argument("--fft", default="fused", choices=["fused", "separate", "rocfft"],
                    help="long-range transforms: fused hand-written kernels (default), the same with the "
                         "interpolation in its own kernel, or rocFFT")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the oracle (CPU baseline and errors)")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 path")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-correction path")
    ap.add_argument("--only", choices=["fp32", "fp64", "learned"],
                    help="time one path only and skip every other measurement (for profiler runs)")
    ap.add_argument("--weights", default=None, help="trained network of the learned correction (.npz; default: "
                                                     "the package's)")
    ap.add_argument("--advance", type=int, default=0, metavar="STEPS",
                    help="time STEPS SSPRK33 steps with the filaments on the device (vpbs_advance, 3 evaluations "
                         "per step; learned evaluations when a network is available) and exit")
    ap.add_argument("--dt", type=float, default=1e-5, help="time step of --advance")
    ap.add_argument("--fp64-advance", action="store_true", help="--advance with --weights none: the plain FP64 path")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB copy kernel (counter calibration for rocprofv3 --pmc) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: the library runs on one GPU "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    vp = ge.load_package()
    t = time.perf_counter()
    fs, spl, p32, p64, pd = build_workload(vp, args.config, args.sigma, args.m, args.fft, host_points=not args.advance)
    log(f"workload: {len(fs)} filaments, Np = {pd.Np}, Nq = {pd.Nq}, Ns = {spl.Ns}, α = {spl.α:.4f}, "
        f"rcut = {spl.rcut:.5f} (host prep {time.perf_counter() - t:.1f} s)")

    if args.advance:
        r = time_advance(vp, pd, spl, args.advance, args.warmup, args.dt, args.weights,
                         np.float64 if args.fp64_advance else np.float32)
        r.update({"workload": f"config {args.config}: {len(fs)} filaments, {pd.Np} nodes, Fourier grid {spl.Ns}",
                  "node_velocities_per_s": pd.Np / (r["ms_per_evaluation"] * 1e-3)})
        print(json.dumps(r), flush=True)
        return
    if args.only:
        if args.only == "learned":
            r = time_learned(vp, pd, spl, args.steps, args.warmup, args.weights)
            log(f"learned path: {r['ms_per_step']:.3f} ms/step; {json.dumps(r['sampled_check'])}" if r
                else "learned path: no weights")
        else:
            ms, st, _, _, _ = time_path(vp, p32 if args.only == "fp32" else p64, pd, args.steps, args.warmup)
            log(f"{args.only} path: {ms:.3f} ms/step; stages {json.dumps({k: round(v, 3) for k, v in st.items()})}")
        return

    ms32, st32, v32, grid, cells = time_path(vp, p32, pd, args.steps, args.warmup)
    log(f"FP32 path: {ms32:.3f} ms/step; stages {json.dumps({k: round(v, 3) for k, v in st32.items()})}")
    res64 = None
    if not args.no_fp64:
        ms64, st64, v64, _, _ = time_path(vp, p64, pd, args.steps, args.warmup)
        log(f"FP64 path: {ms64:.3f} ms/step; stages {json.dumps({k: round(v, 3) for k, v in st64.items()})}")
        res64 = {"value": pd.Np / (ms64 * 1e-3), "ms_per_step": ms64, "dtype": "f64",
                 "stages_ms": {k: round(v, 4) for k, v in st64.items()}}
    learned = None
    if not args.no_learned:
        learned = time_learned(vp, pd, spl, args.steps, args.warmup, args.weights)
        if learned:
            log(f"learned path: {learned['ms_per_step']:.3f} ms/step; {json.dumps(learned['sampled_check'])}; "
                f"stages {json.dumps(learned['stages_ms'])}")

    pdt = time_point_data(vp, p64, pd)
    log(f"point data: {json.dumps(pdt)}")
    e2e = time_end_to_end(vp, p32, pd, reps=2)
    log(f"end to end: {json.dumps(e2e)}")

    # roofline of the dominant long-range kernel of the FP32 path (our own kernels only); the
    # hand-written transforms take the inverse x transform into the interpolation on 192-,
    # 384- and 768-cell x-grids (FP32)
    fused = args.fft == "fused"
    slab = fused and grid[0] in (192, 384, 768)
    lr_kernels = ["spread", "fourier", "interp"]
    dom = max(lr_kernels, key=lambda k: st32.get(k, 0.0))
    nbytes = stage_bytes(dom, 4, grid, spl.Ns, pd.Np, pd.Nq, slab)
    achieved = nbytes / (st32[dom] * 1e-3) / 1e9
    per_stage = {k: round(stage_bytes(k, 4, grid, spl.Ns, pd.Np, pd.Nq, slab) / (st32[k] * 1e-3) / 1e9, 1)
                 for k in ("spread", "fft_forward", "fourier", "fft_backward", "interp") if st32.get(k, 0) > 0}

    kname = {"spread": f"spread_rows_kernel<float, {2 * args.m}>",
             "fourier": f"spec_z_fourier<float, {grid[2]}>" if fused or args.fft == "separate" else "fourier_kernel<float>",
             "interp": f"x_inverse_interp<float, {2 * args.m}, {grid[0]}>" if slab
             else f"interp_kernel_wave<float, {2 * args.m}>"}[dom]
    # the same stages with each range alone (the other stream idle): what the kernels reach by
    # themselves, against the figures above where the two streams share the GPU
    ms_lr, st_lr, _, _, _ = time_path(vp, p32, pd, args.steps, args.warmup, vp._lib.NO_SHORTRANGE)
    ms_sr, st_sr, _, _, _ = time_path(vp, p32, pd, args.steps, args.warmup, vp._lib.NO_LONGRANGE)
    alone = {"long_range_ms": round(ms_lr, 4), "short_range_ms": round(ms_sr, 4),
             "stages_ms": {k: round(v, 4) for k, v in {**st_sr, **st_lr}.items() if v > 0}}
    log(f"alone: {json.dumps(alone)}")
    pair = pair_kernel_report(vp, p32, pd, st32.get("pair", 0.0))
    pair["kernel"] = "pair_wave_kernel<float, true, false, false>"
    log(f"pair kernel: {json.dumps(pair)}")
    default_workload = (args.config, args.sigma, args.m, args.fft) == (3, 1.5, 4, "fused")
    traffic = measured_traffic(kname) if default_workload else None

    result = {
        "metric": METRIC,
        "value": pd.Np / (ms32 * 1e-3),
        "unit": "node-velocities/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"config {args.config}: {len(fs)} random vortex rings, {pd.Np} nodes, {pd.Nq} quadrature "
                        f"points, Fourier grid {spl.Ns[0]}^3 (NUFFT grid {grid[0]}^3, sigma {args.sigma}, m {args.m}), beta 3.5, "
                        f"FP32-bulk path, " + {"fused": "fused hand-written transforms",
                                               "separate": "hand-written transforms, separate interpolation",
                                               "rocfft": "rocFFT transforms"}[args.fft],
            "nodes": pd.Np, "quadrature_points": pd.Nq, "fourier_grid": list(spl.Ns),
            "nufft_grid": list(grid), "nufft_sigma": args.sigma, "nufft_m": args.m, "cells": list(cells),
            "alpha": spl.α, "rcut": spl.rcut, "fft": args.fft,
        },
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": os.path.relpath(PMC_SUMMARY, os.path.dirname(os.path.abspath(__file__)))
            if traffic is not None else None,
            "algorithmic_bytes": nbytes,
            "avg_ms": round(st32[dom], 4),
            "alone": {"avg_ms": round(st_lr[dom], 4),
                      "achieved": round(nbytes / (st_lr[dom] * 1e-3) / 1e9, 1),
                      "frac": round(nbytes / (st_lr[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        },
        "stages_ms": {k: round(v, 4) for k, v in st32.items()},
        "ranges_alone": alone,
        "stage_algorithmic_GBps": per_stage,
        "pair_kernel": pair,
        "point_data": pdt,
        "end_to_end_ms": e2e,
    }
    if res64 is not None:
        result["fp64_path"] = res64
    if learned is not None:
        vl = learned.pop("velocity")
        vl_cheap = learned.pop("velocity_cheap")
        learned["speedup_vs_fp64"] = (res64["ms_per_step"] / learned["ms_per_step"]) if res64 else None
        learned["dtype"] = "f32 (bf16 network)"
        result["learned_correction"] = learned
    if not args.no_cpu_baseline:
        O = ge.load_oracle()
        log("running the FP64 oracle on the whole workload ...")
        ref, cb = oracle_run(O, pd, p64)
        log(f"CPU baseline: {cb['value']:.3e} node-velocities/s ({cb['sample']})")
        result["cpu_baseline"] = cb
        errs = {"fp32": errors_vs_oracle(vp, gpu_fields(vp, p32, pd), ref, pd, p64)}
        if res64 is not None:
            errs["fp64"] = errors_vs_oracle(vp, gpu_fields(vp, p64, pd), ref, pd, p64)
        if learned is not None:
            errs["learned"] = {"velocity": rel_l2(vl, ref["velocity"])}
            errs["cheap_without_network"] = {"velocity": rel_l2(vl_cheap, ref["velocity"])}
            learned["rel_l2_velocity_vs_oracle"] = errs["learned"]["velocity"]
            learned["cheap"]["rel_l2_velocity_vs_oracle"] = errs["cheap_without_network"]["velocity"]
        errs["against"] = "one full FP64 oracle evaluation (oracle/) of the same point data"
        log(f"errors vs oracle: {json.dumps(errs)}")
        result["rel_l2_error"] = errs
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
