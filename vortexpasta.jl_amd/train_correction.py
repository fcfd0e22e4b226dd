#!/usr/bin/env python3
"""Train the learned correction of the β = 2 splitting on the GPU and save its weights.

Training set: seeded random-ring tangles of the two benchmark distributions,
  * radii U[0.5, 1.5], node spacing 2π/256, 1000 rings (≈256k nodes; benchmark/benchmarks.jl
    generator with the 16k-node configuration's parameters), and
  * radii U[0.3, 1.0], node spacing 2π/512, 3000 rings (≈1.0M nodes),
each evaluated with the cheap splitting (β = 2 on a 256³ grid) in FP64 and paired with the exact
converged − cheap difference (β = 4 on the same grid). The network is trained with the library's
HIP forward, backward and Adam kernels and evaluated on held-out tangles of the first
distribution with its BF16-MFMA inference, against the converged splitting on every node.

    python vortexpasta.jl_amd/train_correction.py --out weights.npz [--tangles 40 --big 4 --held-out 20]
"""

from __future__ import annotations

import argparse
import math
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def log(*a):
    print(*a, flush=True)


def main():
    import __graft_entry__ as ge

    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True)
    ap.add_argument("--tangles", type=int, default=40, help="≈256k-node training tangles")
    ap.add_argument("--big", type=int, default=4, help="≈1M-node training tangles")
    ap.add_argument("--held-out", type=int, default=20)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=2e-5)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--Ns", type=int, default=256, help="cheap Fourier grid per direction")
    ap.add_argument("--beta-cheap", type=float, default=2.5, help="β of the cheap splitting (α = kmax/2β)")
    ap.add_argument("--beta-short", type=float, default=None, help="cheap cut-off rcut = beta_short/α")
    ap.add_argument("--feature-factor", type=float, default=2.0, help="feature radius / cheap cut-off")
    ap.add_argument("--nufft-m", type=int, default=4, help="NUFFT kernel half-support of the cheap evaluation")
    args = ap.parse_args()

    vp = ge.load_package()
    L = vp.learned
    two_pi = 2 * math.pi
    st = L.LearnedSettings.create(two_pi, args.Ns, beta_cheap=args.beta_cheap, beta_short=args.beta_short,
                                  feature_factor=args.feature_factor, nufft_m=args.nufft_m)
    log(f"cheap {st.cheap}\nconverged {st.converged}\nfeature radius {st.feature_radius:.5f}")

    def small(seed):
        return vp.tangles.random_rings(1000, 0.5, 1.5, two_pi / 256, seed=seed)

    def big(seed):
        return vp.tangles.random_rings(3000, 0.3, 1.0, two_pi / 512, seed=seed)

    t0 = time.perf_counter()
    sets = [L.training_pairs((small(1000 + k) for k in range(args.tangles)), st, log=log)]
    if args.big:
        sets.append(L.training_pairs((big(2000 + k) for k in range(args.big)), st, log=log))
    data = L.TrainingSet.concatenate(sets) if len(sets) > 1 else sets[0]
    del sets
    log(f"pairs: {len(data)} samples in {time.perf_counter() - t0:.1f} s; target rms {data.stats.t_rms}")

    net = L.CorrectionNetwork.initial(seed=args.seed)
    t0 = time.perf_counter()
    L.fit(net, data, steps=args.steps, batch=args.batch, lr=args.lr, lr_final=args.lr_final, seed=args.seed, log=log)
    log(f"fit: {time.perf_counter() - t0:.1f} s")
    del data

    # held-out evaluation (FP32 bulk, BF16 inference; the check over every node gives the error
    # of the corrected and of the cheap velocity against the converged splitting)
    lb = L.LearnedBiotSavart(st, net, T=np.float32, sample_fraction=1.0, tol=1.0)
    errs = []
    for k in range(args.held_out):
        fs = small(5000 + k)
        _, rep = lb.velocity_on_nodes(fs, fetch=False)
        errs.append((rep["err_uncorrected"], rep["err_corrected"]))
        log(f"held-out {k}: cheap {rep['err_uncorrected']:.3e}, corrected {rep['err_corrected']:.3e}, "
            f"gain {rep['err_uncorrected'] / rep['err_corrected']:.2f}")
    lb.close()
    e = np.array(errs)
    net.meta.update({
        "settings": st.to_meta(),
        "training": {"small_tangles": [1000, 1000 + args.tangles], "big_tangles": [2000, 2000 + args.big]},
        "held_out": {"seeds": [5000, 5000 + args.held_out], "err_cheap": e[:, 0].tolist(),
                     "err_corrected": e[:, 1].tolist(), "mean_gain": float(np.mean(e[:, 0] / e[:, 1]))},
    })
    net.save(args.out)
    log(f"saved {args.out}: held-out mean gain {net.meta['held_out']['mean_gain']:.2f}")


if __name__ == "__main__":
    main()
