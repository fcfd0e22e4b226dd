"""Pins the CPU restatement of the learned correction's features (oracle/learned_oracle.py) by
geometry whose answer is known: straight parallel lines, an isolated ring, and invariance under
a rotation of the whole configuration (the features live in each node's own frame)."""
import importlib.util
import math
import pathlib

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
TWO_PI = 2 * math.pi


@pytest.fixture(scope="module")
def LO():
    spec = importlib.util.spec_from_file_location("learned_oracle", ROOT / "oracle" / "learned_oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _feats(vp, LO, fs, alpha, radius, L=TWO_PI):
    spl = vp.GaussianSplitting.create(Ls=L, β=2.0, Ns=64)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl)
    b = vp.FilamentBatch(fs)
    pd = vp.add_point_charges(fs, p, batch=b)
    return LO.features(pd.nodes, pd.sp, pd.spp, b.offs, b.Ns, (L, L, L), alpha, radius), b


def test_parallel_lines(vp, LO):
    """Two straight lines along z, 0.3 apart: every node sees exactly one strand, at perpendicular
    distance 0.3 along the frame's normal-binormal plane, with a tangent ±t."""
    d, alpha = 0.3, 10.0
    fs = [vp.filaments.init(vp.define_curve(vp.PeriodicLine(), scale=TWO_PI, translate=(3.0, 3.0, math.pi)), 40),
          vp.filaments.init(vp.define_curve(vp.PeriodicLine(), scale=TWO_PI, translate=(3.0 + d, 3.0, math.pi),
                                            orientation=-1), 40)]
    (X, F), _ = _feats(vp, LO, fs, alpha, 0.5)
    assert np.all(X[:, 1] < 1e-8)  # no curvature
    assert np.all(X[:, 3] == 1.0)
    slot0 = X[:, 6:12]
    assert np.allclose(np.linalg.norm(slot0[:, :3], axis=1), alpha * d, rtol=1e-6)
    assert np.allclose(slot0[:, 0], 0.0, atol=1e-6)  # perpendicular to t
    assert np.allclose(slot0[:, 3], -1.0, atol=1e-6)  # antiparallel strand
    assert np.all(X[:, 12:] == 0.0)  # nothing else within the radius


def test_isolated_ring(vp, LO):
    """A ring of radius 0.2 with a feature radius 0.3: the whole ring lies within the node's own
    stretch, so no strand is reported; the curvature entry is 1/(Rα)."""
    R, alpha = 0.2, 20.0
    fs = [vp.filaments.init(vp.define_curve(vp.Ring(), scale=R, translate=math.pi), 48)]
    (X, F), _ = _feats(vp, LO, fs, alpha, 0.3)
    assert np.allclose(X[:, 1], 1 / (R * alpha), rtol=2e-3)
    assert np.all(X[:, 6:] == 0.0)
    # frame: n points to the centre, b along the ring axis (z)
    assert np.allclose(np.abs(F[:, 8]), 1.0, atol=1e-5)


def test_rotation_invariance(vp, LO):
    """Rotating a compact configuration (far from the box faces) leaves every feature unchanged."""
    rng = np.random.default_rng(4)
    alpha, radius = 8.0, 0.5
    centre = np.array([math.pi] * 3)

    def make(rot):
        fs = []
        for k in range(3):
            c = centre + rot @ np.array([0.25 * k, 0.1 * (k % 2), 0.0])
            r2 = rot @ vp.tangles.random_rotation(np.random.default_rng(10 + k))
            fs.append(vp.filaments.init(vp.define_curve(vp.Ring(), scale=0.4, translate=c, rotate=r2), 60))
        return fs

    (X0, _), _ = _feats(vp, LO, make(np.eye(3)), alpha, radius)
    (X1, _), _ = _feats(vp, LO, make(vp.tangles.random_rotation(rng)), alpha, radius)
    assert (np.abs(X0[:, 6:]).sum(axis=1) > 0).mean() > 0.3  # strands are present
    assert np.abs(X1 - X0).max() < 1e-4
