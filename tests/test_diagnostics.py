"""Host diagnostics: kinetic energy from streamfunction values on nodes.

``kinetic_energy_from_streamfunction`` follows ``src/Diagnostics/energy.jl:104-119``.
It is pinned by the vortex-ring closed form of ``test/ring.jl:108-121`` and by two
invariants that a dropped weight or a wrong normalisation would break.
"""
import math

import numpy as np
import pytest


def _ring(vp, N=32, R=math.pi / 3, Γ=2.4, a=1e-6, Δ=0.25, gamma=None):
    f = vp.filaments.init(vp.define_curve(vp.Ring(), scale=R, translate=math.pi), N)
    spl = vp.GaussianSplitting.create(Ls=2 * math.pi, β=3.5, Ns=64)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=Γ, a=a, Δ=Δ, splitting=spl, quadrature=vp.GaussLegendre(4),
                                  lia_segment_fraction=gamma)
    return f, p, vp.add_point_charges([f], p)


def test_energy_ring_closed_form(vp, O):
    """Open-domain ring energy Γ²R/2·(ln(8R/a) − (Δ+1)) to 1e-4 with lia_segment_fraction = 0.4
    (``test/ring.jl:42-121``)."""
    R, Γ, a, Δ = math.pi / 3, 2.4, 1e-6, 0.25
    f, p, pd = _ring(vp, gamma=0.4)
    psi = O.compute_open_domain(pd, p, ("streamfunction",))["streamfunction"]
    E = vp.diagnostics.kinetic_energy_from_streamfunction(pd.batch, psi, Γ)  # no periods: V = 1
    assert E == pytest.approx(Γ ** 2 * R / 2 * (math.log(8 * R / a) - (Δ + 1)), rel=1e-4)
    assert E == pytest.approx(O.energy_from_streamfunction(pd, psi, p, open_domain=True), rel=1e-13)
    # periodic normalisation divides by the cell volume
    Ep = vp.diagnostics.kinetic_energy_from_streamfunction(pd.batch, psi, Γ, p.Ls, sp=pd.sp)
    assert Ep == pytest.approx(E / (2 * math.pi) ** 3, rel=1e-13)


def test_energy_invariants(vp):
    """A uniform ψ integrates to zero on a closed curve; ψ·s′ = 1 gives Γ/2 times the filament period."""
    f, p, pd = _ring(vp, N=48)
    b = pd.batch
    E0 = vp.diagnostics.kinetic_energy_from_streamfunction(b, np.tile([[0.3], [-1.2], [0.7]], (1, b.Np)), 1.0)
    assert abs(E0) < 1e-13
    # ψ = s′/|s′|² makes ψ·s′ = 1, so the weights (t_{i+1} − t_{i−1})/2 must sum to the period T
    psi1 = pd.sp / np.sum(pd.sp ** 2, axis=0)
    E1 = vp.diagnostics.kinetic_energy_from_streamfunction(b, psi1, 2.0)
    assert E1 == pytest.approx(2.0 / 2 * b.T[0], rel=1e-13)
    # and ψ = s′ weighs |s′|² = (L/T)² on the ring (uniform nodes), L the arc length 2πR
    E2 = vp.diagnostics.kinetic_energy_from_streamfunction(b, pd.sp, 2.0)
    assert E2 == pytest.approx((2 * math.pi ** 2 / 3) ** 2 / b.T[0], rel=1e-6)
    with pytest.raises(ValueError):
        vp.diagnostics.kinetic_energy_from_streamfunction(b, np.zeros((2, b.Np)), 1.0)
