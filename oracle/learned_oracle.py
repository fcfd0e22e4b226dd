"""CPU restatement of the learned correction's per-node features (numpy, brute force).

TEST INFRASTRUCTURE ONLY: the checker of ``feature_kernel`` (``csrc/learned.hip``)
on small inputs. The features have no counterpart in the reference; they are
defined by ``include/vpbs.h`` (vpbs_learned_features):

* frame of node i: t = s′/|s′|, κ = (s″ − (s″·t)t)/|s′|², n = κ/|κ|, b = t × n
  (|κ| ≤ 1e-9·α: n = t × e/|t × e|, e the axis least aligned with t);
* row[1] = |κ|/α, row[3] = 1 (curvature vector and tangent in the frame);
* strands: nodes j ≠ i within the radius R (minimum image, r² = (x² + y²) + z²
  in FP64 with r = s_j − x), outside i's own stretch (the runs of consecutive
  neighbours of i along its filament that lie within R, at most 32 each way),
  where the distance to i has a local minimum along j's filament. In FP32 with
  r' = float(r), w∓ = float(s_{j∓1} − s_j) (minimum image):
  2 r'·w⁻ + |w⁻|² ≥ 0 and 2 r'·w⁺ + |w⁺|² > 0 (that is d²(j−1) ≥ d²(j) < d²(j+1)).
  For each, with t_j = float(s′_j/|s′_j|): r_p = r' − (r'·t_j)t_j (FP32); the 16
  with the smallest |r_p|² (FP32; ties: smaller j) fill slots 6 + 6k .. 6 + 6k + 5
  with (α r_p, t_j) projected on i's frame in FP64.
  Dot products are summed as (a₀b₀ + a₁b₁) + a₂b₂ without fused multiply-adds.

The check against the GPU is to FP32 rounding of the stored values.
"""

from __future__ import annotations

import numpy as np


def frames(sp, spp, alpha):
    """(t, n, b, |κ|) of every node; sp, spp (3, Np)."""
    s1, s2 = np.asarray(sp, np.float64), np.asarray(spp, np.float64)
    m2 = np.sum(s1 * s1, axis=0)
    t = s1 / np.sqrt(m2)
    st = np.sum(s2 * t, axis=0)
    k = (s2 - st * t) / m2
    kn = np.sqrt(np.sum(k * k, axis=0))
    n = np.empty_like(t)
    big = kn > 1e-9 * alpha
    n[:, big] = k[:, big] / kn[big]
    for i in np.nonzero(~big)[0]:
        a = np.abs(t[:, i])
        e = 0 if (a[0] <= a[1] and a[0] <= a[2]) else (1 if a[1] <= a[2] else 2)
        u = np.zeros(3)
        u[e] = 1.0
        c = np.cross(t[:, i], u)
        n[:, i] = c / np.linalg.norm(c)
    b = np.cross(t.T, n.T).T
    return t, n, b, kn


def _sep(x, y, L):
    r = y - x
    Lh = L / 2
    r = np.where(r >= Lh, r - L, r)
    r = np.where(r < -Lh, r + L, r)
    return r


def features(nodes, sp, spp, offs, counts, Ls, alpha, radius, K=16, n_in=128):
    """Feature rows (Np, n_in) and frames (Np, 9) of folded nodes (3, Np); filament f holds
    nodes offs[f] .. offs[f] + counts[f] - 1."""
    nodes = np.asarray(nodes, np.float64)
    Np = nodes.shape[1]
    L = np.asarray(Ls, np.float64)[:, None]
    t, n, b, kn = frames(sp, spp, alpha)
    fid = np.repeat(np.arange(len(counts)), counts)
    off = np.asarray(offs)[fid]
    N = np.asarray(counts)[fid]
    pos = np.arange(Np) - off
    prev = off + (pos + N - 1) % N
    nxt = off + (pos + 1) % N
    tu = (np.asarray(sp) / np.sqrt(np.sum(np.asarray(sp, np.float64) ** 2, axis=0))).astype(np.float32)
    wm = _sep(nodes, nodes[:, prev], L).astype(np.float32)  # s_{j-1} - s_j
    wp = _sep(nodes, nodes[:, nxt], L).astype(np.float32)   # s_{j+1} - s_j
    f32 = np.float32

    def dot(a, b):  # (a0 b0 + a1 b1) + a2 b2, rounded at each step in the arrays' type
        return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]

    R2 = radius * radius
    X = np.zeros((Np, n_in))
    F = np.zeros((Np, 9))
    for i in range(Np):
        r = _sep(nodes[:, i : i + 1], nodes, L)  # s_j - x_i
        d2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]
        Ni, pi, oi = N[i], pos[i], off[i]
        # own stretch
        dp = dm = 32
        for k in range(1, 33):
            if d2[oi + (pi + k) % Ni] > R2:
                dp = k - 1
                break
        for k in range(1, 33):
            if d2[oi + ((pi - k) % Ni + Ni) % Ni] > R2:
                dm = k - 1
                break
        ok = d2 <= R2
        ok[i] = False
        same = fid == fid[i]
        fo = (pos - pi) % Ni
        ok &= ~(same & ((fo <= dp) | (Ni - fo <= dm)))
        rf = r.astype(np.float32)
        gm = f32(2) * dot(rf, wm) + dot(wm, wm)
        gp = f32(2) * dot(rf, wp) + dot(wp, wp)
        ok &= (gm >= 0) & (gp > 0)
        js = np.nonzero(ok)[0]
        rj = rf[:, js]
        tj = tu[:, js]
        rt = dot(rj, tj)
        rp = rj - rt * tj
        key = dot(rp, rp)
        order = np.lexsort((js, key))[:K]
        E = np.stack([t[:, i], n[:, i], b[:, i]])  # rows: frame vectors (FP64)
        X[i, 1] = kn[i] / alpha
        X[i, 3] = 1.0
        for slot, q in enumerate(order):
            pq = rp[:, q].astype(np.float64)
            tq = tj[:, q].astype(np.float64)
            for e in range(3):
                X[i, 6 + 6 * slot + e] = alpha * ((pq[0] * E[e, 0] + pq[1] * E[e, 1]) + pq[2] * E[e, 2])
                X[i, 9 + 6 * slot + e] = (tq[0] * E[e, 0] + tq[1] * E[e, 1]) + tq[2] * E[e, 2]
        F[i] = np.concatenate([t[:, i], n[:, i], b[:, i]])
    return X.astype(np.float32), F.astype(np.float32)
