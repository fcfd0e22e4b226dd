"""Time stepping (SSPRK33) host logic, CPU only."""
import math

import numpy as np


def test_ssprk33_linear_ode(vp):
    """dx/dt = λx: one step multiplies by 1 + z + z²/2 + z³/6 (z = λΔt), the stability
    polynomial of the three-stage third-order SSP scheme (SSPRK33.jl:17-39)."""
    lam = np.array([-0.7, 0.3, -2.1])
    x0 = np.array([1.0, -0.5, 2.0])
    dt = 0.13
    x1 = vp.timestepping.ssprk33_update(x0, lambda x, stage: lam * x, dt)
    z = lam * dt
    assert np.allclose(x1, x0 * (1 + z + z * z / 2 + z ** 3 / 6), rtol=1e-15, atol=0)


def test_ssprk33_third_order(vp):
    """Global error of dx/dt = cos(t)-like autonomous system x' = (-x1, x0) falls as Δt³."""
    def run(n):
        x = np.array([1.0, 0.0])
        for _ in range(n):
            x = vp.timestepping.ssprk33_update(x, lambda y, s: np.array([-y[1], y[0]]), 1.0 / n)
        return np.linalg.norm(x - np.array([math.cos(1.0), math.sin(1.0)]))
    e1, e2 = run(20), run(40)
    assert 7.0 < e1 / e2 < 9.0  # 2³ = 8


def test_stage_batches_keep_knots(vp):
    """Intermediate stages keep the knots; the final state recomputes them (advect.jl:20-72)."""
    fs = vp.tangles.random_rings(3, 0.5, 1.0, 2 * math.pi / 32, seed=4)
    b = vp.FilamentBatch(fs)
    X = b.X + 0.01 * np.sin(b.X)
    kept = b.with_nodes(X, keep_knots=True)
    new = b.with_nodes(X, keep_knots=False)
    assert np.array_equal(kept.ts, b.ts) and np.array_equal(kept.T, b.T)
    assert not np.array_equal(new.ts, b.ts)
    # the interpolant still passes through the nodes with kept knots
    s = kept.evaluate(kept.fid, kept.local, np.zeros(kept.Np), 0)
    assert np.abs(s - X).max() < 1e-13
