"""Filament geometry: quintic spline interpolation, derivatives, point data layout."""
import math

import numpy as np
import pytest


def _ring(vp, N, R=math.pi / 3):
    return vp.filaments.init(vp.define_curve(vp.Ring(), scale=R, translate=math.pi), N)


def test_interpolates_nodes_and_is_periodic(vp):
    f = _ring(vp, 40)
    b = vp.FilamentBatch([f])
    x0 = b.evaluate(b.fid, b.local, np.zeros(b.Np))
    x1 = b.evaluate(b.fid, b.local, np.ones(b.Np))
    assert np.abs(x0 - b.X).max() < 1e-14
    assert np.abs(x1 - b.X[(b.local + 1) % 40]).max() < 1e-14


def test_derivatives_match_finite_differences(vp):
    f = vp.filaments.init(vp.define_curve(vp.TrefoilKnot(), scale=math.pi / 3, translate=math.pi / 3), 48)
    b = vp.FilamentBatch([f])
    z = np.full(b.Np, 0.37)
    h = 1e-6
    dt = b.segment_dt()[:, None]
    for k in (1, 2):
        fd = (b.evaluate(b.fid, b.local, z + h, k - 1) - b.evaluate(b.fid, b.local, z - h, k - 1)) / (2 * h * dt)
        an = b.evaluate(b.fid, b.local, z, k)
        assert np.abs(fd - an).max() / np.abs(an).max() < 1e-7


@pytest.mark.parametrize("N,tol", [(32, 3e-6), (64, 2e-7)])
def test_ring_curvature_and_length(vp, N, tol):
    R = math.pi / 3
    b = vp.FilamentBatch([_ring(vp, N, R)])
    s1 = b.evaluate(b.fid, b.local, np.zeros(b.Np), 1)
    s2 = b.evaluate(b.fid, b.local, np.zeros(b.Np), 2)
    kappa = np.linalg.norm(np.cross(s1, s2), axis=1) / np.linalg.norm(s1, axis=1) ** 3
    assert np.abs(kappa * R - 1).max() < tol


def test_trefoil_curvature(vp):
    """Quintic-spline curvature of the trefoil at N = 96 (test/trefoil.jl:50-86: rms < 1e-4, max < 2e-4)."""
    from numpy import pi, sin, cos

    S = vp.define_curve(vp.TrefoilKnot(), scale=pi / 3, translate=pi / 3)
    N = 96
    f = vp.ClosedFilament(S(np.arange(N) / N), np.zeros(3))
    b = vp.FilamentBatch([f])
    s1 = b.evaluate(b.fid, b.local, np.zeros(N), 1)
    s2 = b.evaluate(b.fid, b.local, np.zeros(N), 2)
    k_num = np.linalg.norm(np.cross(s1, s2), axis=1) / np.linalg.norm(s1, axis=1) ** 3
    # analytic curvature of the trefoil (u = 2πt)
    u = 2 * pi * np.arange(N) / N
    d1 = np.stack([cos(u) + 4 * cos(2 * u), -sin(u) + 4 * sin(2 * u), -3 * cos(3 * u)], 1)
    d2 = np.stack([-sin(u) - 8 * sin(2 * u), -cos(u) + 8 * cos(2 * u), 9 * sin(3 * u)], 1)
    k_ex = np.linalg.norm(np.cross(d1, d2), axis=1) / np.linalg.norm(d1, axis=1) ** 3 / (pi / 3)
    kmax = np.abs(k_ex).max()
    err = np.sqrt(np.mean((k_num - k_ex) ** 2)) / kmax
    err_max = np.abs(k_num - k_ex).max() / kmax
    assert err < 1e-4 and err_max < 2e-4  # same metric and bounds as the reference test


def test_point_data_layout(vp):
    fs = [_ring(vp, 16), _ring(vp, 20, R=0.5)]
    spl = vp.GaussianSplitting.create(Ls=2 * math.pi, β=3.5, Ns=64)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl)
    pd = vp.add_point_charges(fs, p)
    assert pd.Np == 36 and pd.Nq == 108 and pd.nquad == 3
    assert pd.node_idx_prev[0] == 15 and pd.node_idx_prev[16] == 35 and pd.node_idx_prev[17] == 16
    assert (pd.points >= 0).all() and (pd.points < 2 * math.pi).all()
    # charges of a closed filament sum to zero; sub-segment lengths sum to the length
    assert np.abs(pd.charges[:, :48].sum(axis=1)).max() < 1e-13
    assert pd.dlp[:16].sum() == pytest.approx(2 * math.pi * math.pi / 3, rel=3e-7)
    assert np.allclose(pd.dlm[1:16], pd.dlp[:15])


def test_fold_coordinates(vp):
    x = np.array([-1e-20, 0.0, 2 * math.pi, -2 * math.pi - 1.0, 13.0])
    y = vp.filaments.fold_coordinates_periodic(x, 2 * math.pi)
    assert (y >= 0).all() and (y < 2 * math.pi).all()
    assert y[0] == 0.0  # tiny negative folds to 0, not L (min_distance.jl:6-11)


def test_unclosed_line_offset(vp):
    S = vp.define_curve(vp.PeriodicLine(x=lambda t: 0.1 * np.sin(2 * np.pi * t)), scale=2 * math.pi)
    f = vp.filaments.init(S, 24)
    assert f.offset[2] == pytest.approx(2 * math.pi)
    b = vp.FilamentBatch([f])
    x1 = b.evaluate(b.fid, b.local, np.ones(24))
    nxt = (b.local + 1) % 24
    expect = b.X[nxt] + np.where((b.local == 23)[:, None], f.offset, 0)
    assert np.abs(x1 - expect).max() < 1e-13
