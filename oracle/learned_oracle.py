"""CPU restatement of the learned correction's per-node features (numpy, brute force).

TEST INFRASTRUCTURE ONLY: the checker of ``feature_kernel`` (``csrc/learned.hip``)
on small inputs. The features have no counterpart in the reference; they are
defined by ``include/vpbs.h`` (vpbs_learned_features):

* frame of node i: t = s′/|s′|, κ = (s″ − (s″·t)t)/|s′|², n = κ/|κ|, b = t × n
  (|κ| ≤ 1e-9·α: n = t × e/|t × e|, e the axis least aligned with t);
* row[1] = |κ|/α, row[3] = 1 (curvature vector and tangent in the frame);
* strands: nodes j ≠ i within the radius R (minimum image, r² = (x² + y²) + z²),
  outside i's own stretch (the runs of consecutive neighbours of i along its
  filament that lie within R, at most 32 each way), where the distance to i has
  a local minimum along j's filament (r²_j ≤ r²_prev, r²_j < r²_next); for each,
  r_p = r − (r·t_j)t_j; the 16 with the smallest |r_p|² (ties: smaller j) fill
  slots 6 + 6k .. 6 + 6k + 5 with (α r_p, t_j) in i's frame.

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
    tu = sp / np.sqrt(np.sum(np.asarray(sp) ** 2, axis=0))
    R2 = radius * radius
    X = np.zeros((Np, n_in))
    F = np.zeros((Np, 9))
    for i in range(Np):
        r = _sep(nodes[:, i : i + 1], nodes, L)
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
        ok &= (d2 <= d2[prev]) & (d2 < d2[nxt])
        js = np.nonzero(ok)[0]
        rv = r[:, js]
        tj = tu[:, js]
        rp = rv - np.sum(rv * tj, axis=0) * tj
        key = np.sum(rp * rp, axis=0).astype(np.float32)
        order = np.lexsort((js, key))[:K]
        E = np.stack([t[:, i], n[:, i], b[:, i]])  # rows: frame vectors
        X[i, 1] = kn[i] / alpha
        X[i, 3] = 1.0
        for slot, q in enumerate(order):
            X[i, 6 + 6 * slot : 9 + 6 * slot] = alpha * (E @ rp[:, q])
            X[i, 9 + 6 * slot : 12 + 6 * slot] = E @ tj[:, q]
        F[i] = np.concatenate([t[:, i], n[:, i], b[:, i]])
    return X.astype(np.float32), F.astype(np.float32)
