"""Pins the FP64 CPU oracle against closed forms, brute force and invariances.

Each check follows a reference test (cited) and is chosen so that a plausible
mistake (dropped term, wrong sign, wrong normalisation) fails it; the
planted-mistake tests at the end show that explicitly by breaking the oracle
on purpose and checking that the invariance test notices.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2

TWO_PI = 2 * math.pi


def ring_setup(vp, N, quad=4, R=math.pi / 3, Γ=2.4, a=1e-6, Δ=0.25):
    f = vp.filaments.init(vp.define_curve(vp.Ring(), scale=R, translate=math.pi), N)
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, β=3.5, Ns=64)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=Γ, a=a, Δ=Δ, splitting=spl, quadrature=vp.GaussLegendre(quad))
    return [f], p, vp.add_point_charges([f], p)


def ring_errors(vp, O, N):
    R, Γ, a, Δ = math.pi / 3, 2.4, 1e-6, 0.25
    fs, p, pd = ring_setup(vp, N)
    out = O.compute_open_domain(pd, p, ("velocity", "streamfunction"))
    v, psi = out["velocity"], out["streamfunction"]
    U = np.linalg.norm(v.mean(axis=1))
    U_ex = Γ / (4 * math.pi * R) * (math.log(8 * R / a) - Δ)
    t = pd.sp / np.linalg.norm(pd.sp, axis=0)
    ps = (psi * t).sum(axis=0).mean()
    ps_ex = Γ / (2 * math.pi) * (math.log(8 * R / a) - (1 + Δ))
    E = O.energy_from_streamfunction(pd, psi, p, open_domain=True)
    E_ex = (Γ ** 2 * R / 2) * (math.log(8 * R / a) - (Δ + 1))
    perp = np.linalg.norm(psi - (psi * t).sum(0) * t, axis=0).max() / abs(ps)
    return abs(U - U_ex) / U_ex, abs(ps - ps_ex) / ps_ex, abs(E - E_ex) / E_ex, perp, v


def test_ring_closed_forms(vp, O):
    """Vortex ring velocity / streamfunction / energy (test/ring.jl:35-40,64-121)."""
    eU32, eP32, eE32, perp, v = ring_errors(vp, O, 32)
    assert eU32 < 1e-4  # ring.jl:86 (quintic, GL(4), N = 32)
    assert perp < 1e-10  # ψ parallel to the tangent (ring.jl:101-107)
    # all nodes move with the same velocity along the ring axis
    assert np.abs(v[0]).max() < 1e-12 and np.abs(v[1]).max() < 1e-12
    assert np.std(v[2]) / np.mean(v[2]) < 1e-12
    # Without lia_segment_fraction the local-segment error of ψ and E is O(δ²):
    eU64, eP64, eE64, _, _ = ring_errors(vp, O, 64)
    assert eP64 < 2e-4 and eE64 < 2e-4
    assert 3.5 < eP32 / eP64 < 4.5 and 3.5 < eE32 / eE64 < 4.5 and 3.5 < eU32 / eU64 < 4.5


def test_lia_decomposition(vp, O):
    """LIA + non-LIA = total (test/ring.jl:134-140), periodic Ewald path."""
    fs = vp.tangles.random_rings(6, 0.5, 1.0, TWO_PI / 64, seed=3)
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, β=3.5, Ns=48)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl)
    pd = vp.add_point_charges(fs, p)
    q = ("velocity", "streamfunction")
    tot = O.compute(pd, p, q)
    only = O.compute(pd, p, q, LIA="only")
    off = O.compute(pd, p, q, LIA=False)
    for k in q:
        assert rel_l2(only[k] + off[k], tot[k]) < 1e-14


def _numpy_bruteforce_stats(O, pd, p):
    """Independent numpy enumeration of accepted pairs (r² ≤ rcut², min-image)."""
    L = np.array(p.Ls)[:, None]
    rc2 = p.rcut * p.rcut
    counts = np.zeros(pd.Np, dtype=np.int64)
    hashes = np.zeros(pd.Np, dtype=np.uint64)
    H = np.array([O.lib().or_pair_hash(j) for j in range(pd.Nq)], dtype=np.uint64)
    for i in range(pd.Np):
        r = pd.nodes[:, i : i + 1] - pd.points
        r = np.where(r >= L / 2, r - L, r)
        r = np.where(r < -L / 2, r + L, r)
        r2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]
        acc = r2 <= rc2
        counts[i] = acc.sum()
        with np.errstate(over="ignore"):
            hashes[i] = np.sum(H[acc], dtype=np.uint64)
    return counts, hashes


def test_neighbour_sets_celllists_equal_bruteforce(vp, O):
    """Cell lists find exactly the brute-force pair set (test/cell_lists.jl:43-60,218-239)."""
    fs = vp.tangles.random_rings(24, 0.3, 0.8, TWO_PI / 96, seed=11)
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, β=3.5, Ns=48)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl)
    pd = vp.add_point_charges(fs, p)
    cl = O.compute(pd, p, ("velocity",), longrange_on=False, neighbour_stats=True)["stats"]
    bf = O.neighbour_pairs_bruteforce(pd, p)
    npbf = _numpy_bruteforce_stats(O, pd, p)
    assert np.array_equal(cl[0], bf[0]) and np.array_equal(cl[1], bf[1])
    assert np.array_equal(bf[0], npbf[0]) and np.array_equal(bf[1], npbf[1])
    assert cl[0].sum() > 10 * pd.Np  # the test is not vacuous
    # sums agree to roundoff (cell_lists.jl: rtol 1e-13)
    a = O.compute(pd, p, ("velocity",), longrange_on=False, use_celllist=True)["velocity"]
    b = O.compute(pd, p, ("velocity",), longrange_on=False, use_celllist=False)["velocity"]
    assert rel_l2(a, b) < 1e-13


def trefoil_params(vp, rcut=None, α=None, rtol=1e-6, **kws):
    Ls = (1.5 * math.pi, 1.5 * math.pi, 2 * math.pi)  # test/trefoil.jl:395
    if rcut is not None:
        spl = vp.GaussianSplitting.create(Ls=Ls, rcut=rcut, rtol=rtol)
    else:
        spl = vp.GaussianSplitting.create(Ls=Ls, α=α, rtol=rtol)
    return vp.ParamsBiotSavart.build(np.float64, Γ=7.4, a=1e-5, splitting=spl, **kws)


def trefoil_fs(vp, N=30):
    R = math.pi / 3
    S = vp.define_curve(vp.TrefoilKnot(), scale=R, translate=R)
    return [vp.ClosedFilament(S(np.arange(N) / N), np.zeros(3))]


def test_exact_sum_against_numpy(vp, O):
    """Type-1 exact sum and retained-mode set against a direct numpy evaluation."""
    fs = trefoil_fs(vp, 12)
    p = trefoil_params(vp, α=3.0)
    pd = vp.add_point_charges(fs, p)
    p2 = vp.ParamsBiotSavart.build(np.float64, Γ=7.4, a=1e-5, splitting=vp.GaussianSplitting.create(
        Ls=p.Ls, α=p.α, rcut=p.rcut, Ns=(10, 9, 12)))
    Ns = np.array(p2.Ns, dtype=np.int32)
    Ls = np.array(p2.Ls)
    Nkx = Ns[0] // 2 + 1
    uh = np.zeros((3, Ns[2], Ns[1], Nkx, 2))
    pref = p2.Γ / np.prod(Ls)
    O.lib().or_exact_type1(pd.Nq, O._p(O._c(pd.points)), O._p(O._c(pd.charges)), O._p(Ls), O._p(Ns, O._ip), pref,
                           O._p(uh))
    uh = uh[..., 0] + 1j * uh[..., 1]
    kxm, kym, kzm = O.mode_limits(p2.Ns)
    assert (kxm, kym, kzm) == (4, 4, 5)  # N even: N/2-1 ; N odd: (N-1)/2 (longrange.jl:556-583)
    for kx, ky, kz in [(0, 0, 0), (1, -2, 3), (4, 4, -5), (2, -4, 5)]:
        k = 2 * np.pi * np.array([kx, ky, kz]) / Ls
        want = pref * (pd.charges * np.exp(-1j * (k @ pd.points))).sum(axis=1)
        got = uh[:, kz % Ns[2], ky % Ns[1], kx]
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max() + 1e-15
    # dropped modes are zero: kx = N/2 (r2c Nyquist), ky = -N/2 absent for odd N, kz = -6 (N even)
    assert np.all(uh[:, :, :, 5] == 0) and np.all(uh[:, 6, :, :] == 0)


def test_nufft_vs_exact_sum(vp, O):
    """NUFFT long range agrees with the exact Fourier sum (test/trefoil.jl:135-245,402-406: 1e-5)."""
    fs = trefoil_fs(vp)
    p = trefoil_params(vp, α=3.0)
    pd = vp.add_point_charges(fs, p)
    q = ("velocity", "streamfunction")
    ex = O.longrange(pd, p, q, method="exact")
    nu = O.longrange(pd, p, q, method="nufft")
    for k in q:
        err = np.abs(nu[k] - ex[k]).max() / np.abs(ex[k]).max()
        assert err < 1e-5, (k, err)


def _splitting_maxdiffs(vp, O, fs, rcuts, **kws):
    fields = []
    for rc in rcuts:
        p = trefoil_params(vp, rcut=rc, **kws)
        pd = vp.add_point_charges(fs, p)
        fields.append(O.compute(pd, p, ("velocity", "streamfunction")))
    last = fields[-1]
    dv = max(np.max(np.linalg.norm(f["velocity"] - last["velocity"], axis=0) / np.linalg.norm(last["velocity"], axis=0))
             for f in fields)
    dp = max(np.max(np.linalg.norm(f["streamfunction"] - last["streamfunction"], axis=0)
                    / np.linalg.norm(last["streamfunction"], axis=0)) for f in fields)
    return dv, dp


@pytest.mark.parametrize("avoid_explicit_erf", [True, False])
def test_independence_on_splitting_parameter(vp, O, avoid_explicit_erf):
    """Total v and ψ do not depend on α / rcut (test/trefoil.jl:297-329: vel < 10·rtol, ψ < rtol/5)."""
    rtol = 1e-6
    dv, dp = _splitting_maxdiffs(vp, O, trefoil_fs(vp), [0.3, 0.5, 0.7, 0.9], rtol=rtol,
                                 avoid_explicit_erf=avoid_explicit_erf)
    assert dv < 10 * rtol, dv
    assert dp < rtol / 5, dp


# ----------------------------------------------------------------------------
# Planted mistakes: each breaks the oracle in a plausible way and the checks
# above must notice.


def test_planted_mistake_missing_ewald_factor(vp, O, monkeypatch):
    orig = O._fourier_operator

    def no_ewald(kx, ky, kz, Ls, alpha, quantity, trunc):
        scal, kv = orig(kx, ky, kz, Ls, 1e9, quantity, trunc)  # exp(-k²/4α²) → 1
        return scal, kv

    monkeypatch.setattr(O, "_fourier_operator", no_ewald)
    dv, dp = _splitting_maxdiffs(vp, O, trefoil_fs(vp), [0.3, 0.9])
    assert dv > 1e-4 and dp > 1e-4


def test_planted_mistake_curl_sign(vp, O, monkeypatch):
    orig = O._apply_operator
    monkeypatch.setattr(O, "_apply_operator", lambda w, s, k, q: -orig(w, s, k, q) if q == "velocity" else orig(w, s, k, q))
    dv, _ = _splitting_maxdiffs(vp, O, trefoil_fs(vp), [0.3, 0.9])
    assert dv > 1e-3


def periodic_ring_velocity_error(vp, O, N=64, R=0.5):
    """|U - U_closed_form| / U for a small ring in the periodic box (Ewald path).

    The periodic images shift U by O((R/L)³); for R = 0.5 in a 2π box that is
    below 2e-3, far below the effect of any missing local term.
    """
    Γ, a, Δ = 2.4, 1e-6, 0.25
    fs, p, pd = ring_setup(vp, N, quad=4, R=R)
    v = O.compute(pd, p, ("velocity",))["velocity"]
    U_ex = Γ / (4 * math.pi * R) * (math.log(8 * R / a) - Δ)
    return abs(v[2].mean() - U_ex) / U_ex


def test_periodic_ring_velocity(vp, O):
    assert periodic_ring_velocity_error(vp, O) < 2e-3


def test_planted_mistake_no_self_interaction(vp, O, monkeypatch):
    L = O.lib()

    class NoSelf:
        def __getattr__(self, name):
            if name == "or_self_interaction":
                return lambda *a: None
            return getattr(L, name)

    monkeypatch.setattr(O, "lib", lambda: NoSelf())
    # invisible to the splitting-independence test (the double-counted local
    # integral does not depend on α) ...
    dv, _ = _splitting_maxdiffs(vp, O, trefoil_fs(vp), [0.3, 0.9])
    assert dv < 1e-4
    # ... but caught by the closed form
    assert periodic_ring_velocity_error(vp, O) > 0.1


def test_planted_mistake_lia_cutoff(vp, O, monkeypatch):
    """Using exp(Δ-1)a/2 for the velocity cut-off too (a plausible slip) breaks the ring velocity."""
    monkeypatch.setattr(O, "local_cutoffs", lambda params, ic=None: (math.exp(params.Δ - 1) * params.a / 2,) * 2)
    assert periodic_ring_velocity_error(vp, O) > 1e-2


def _ring_open(vp, O, N, quad, gamma, qns=None, Γ=2.4, a=1e-6, Δ=0.25, R=math.pi / 3):
    f = vp.filaments.init(vp.define_curve(vp.Ring(), scale=R, translate=math.pi), N)
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, β=3.5, Ns=64)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=Γ, a=a, Δ=Δ, splitting=spl, quadrature=vp.GaussLegendre(quad),
                                  quadrature_near_singularity=qns, lia_segment_fraction=gamma)
    pd = vp.add_point_charges([f], p)
    return p, pd, O.compute_open_domain(pd, p, ("velocity", "streamfunction"))


def test_ring_closed_forms_lia_segment_fraction(vp, O):
    """test/ring.jl:42-121 with lia_segment_fraction = 0.4, GL(4), N = 32: U, ψ, E within 1e-4."""
    R, Γ, a, Δ = math.pi / 3, 2.4, 1e-6, 0.25
    p, pd, out = _ring_open(vp, O, 32, 4, 0.4)
    v, psi = out["velocity"], out["streamfunction"]
    U = np.linalg.norm(v.mean(axis=1))
    assert U == pytest.approx(Γ / (4 * math.pi * R) * (math.log(8 * R / a) - Δ), rel=1e-4)
    t = pd.sp / np.linalg.norm(pd.sp, axis=0)
    assert (psi * t).sum(0).mean() == pytest.approx(Γ / (2 * math.pi) * (math.log(8 * R / a) - (1 + Δ)), rel=1e-4)
    E = O.energy_from_streamfunction(pd, psi, p, open_domain=True)
    assert E == pytest.approx((Γ ** 2 * R / 2) * (math.log(8 * R / a) - (Δ + 1)), rel=1e-4)


def test_lia_segment_fraction_convergence(vp, O):
    """test/ring.jl:219-260: γ = 0.2, GL(8) near the singularity, N = 64: σ(v_z)/v_z < 1e-12, error < 2e-6."""
    R, Γ, a, Δ = math.pi / 3, 1.0, 1e-8, 0.25
    p, pd, out = _ring_open(vp, O, 64, 3, 0.2, qns=vp.GaussLegendre(8), Γ=Γ, a=a)
    vz = out["velocity"][2]
    assert np.std(vz) / np.mean(vz) < 1e-12
    U = Γ / (4 * math.pi * R) * (math.log(8 * R / a) - Δ)
    assert abs(np.mean(vz) - U) / U < 2e-6


# --------------------------------------------------------------------------- #
# Long-range field in Fourier space and its consumers (uhat, energy spectrum, band field)


def _ring_cfg1(vp):
    fs = vp.tangles.single_ring()
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, β=3.5, Ns=64)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl)
    return fs, p, vp.add_point_charges(fs, p)


@pytest.mark.parametrize("quantity", ["vorticity", "streamfunction", "velocity"])
def test_uhat_nufft_vs_exact_sum(vp, O, quantity):
    """Unsmoothed Fourier field: NUFFT type 1 against the exact sum (NUFFT tolerance, trefoil.jl:402-406)."""
    fs, p, pd = _ring_cfg1(vp)
    u = O.longrange_fourier(pd, p, quantity)
    ue = O.longrange_fourier(pd, p, quantity, method="exact")
    assert rel_l2(u, ue) < 2e-6
    # Nyquist planes and the k = 0 mode are zero (longrange.jl:556-583; k = 0 dropped by 1/k²)
    assert np.all(u[:, 32] == 0) and np.all(u[:, :, 32] == 0) and np.all(u[..., 32] == 0)
    if quantity != "vorticity":
        assert np.all(u[:, 0, 0, 0] == 0)


def test_uhat_velocity_is_curl_of_streamfunction(vp, O):
    """v̂ = i k × ψ̂ (longrange.jl:160-172), checked against numpy gradients of the real fields."""
    fs, p, pd = _ring_cfg1(vp)
    uv = O.longrange_fourier(pd, p, "velocity")
    up = O.longrange_fourier(pd, p, "streamfunction")
    g = 64
    psi = np.fft.irfftn(up, s=(g, g, g), axes=(1, 2, 3), norm="forward")  # [c][z][y][x]
    v = np.fft.irfftn(uv, s=(g, g, g), axes=(1, 2, 3), norm="forward")
    # spectral derivative of psi along each axis, then curl
    kx, ky, kz = O._wavenumbers((g, g, g), np.array([TWO_PI] * 3))
    dpsi = {}
    for ax, k in ((3, kx[None, None, :]), (2, ky[None, :, None]), (1, kz[:, None, None])):
        dpsi[ax] = np.fft.irfftn(1j * k[None] * up, s=(g, g, g), axes=(1, 2, 3), norm="forward")
    dx, dy, dz = dpsi[3], dpsi[2], dpsi[1]
    curl = np.stack([dy[2] - dz[1], dz[0] - dx[2], dx[1] - dy[0]])
    assert rel_l2(v, curl) < 1e-12


def test_energy_spectrum_parseval(vp, O):
    """Σ E(k) Δk equals the mean kinetic energy ½⟨|v|²⟩ of the real field (numpy irfftn), which pins the
    Hermitian factor 1/2 on the kx = 0 plane and the 1/Δk normalisation (spectra.jl:151-196)."""
    fs, p, pd = _ring_cfg1(vp)
    u = O.longrange_fourier(pd, p, "velocity")
    E = O.energy_spectrum(u, p.Ls, 60, 1.0)  # 60 bins reach past |k| = √3·31
    g = np.fft.irfftn(u, s=(64, 64, 64), axes=(1, 2, 3), norm="forward")
    assert abs(E.sum() - 0.5 * np.mean(np.sum(g ** 2, axis=0))) < 1e-12 * E.sum()
    # same spectrum from the vorticity (|ω̂|²/k²)
    w = O.longrange_fourier(pd, p, "vorticity")
    Ew = O.energy_spectrum(w, p.Ls, 60, 1.0, quantity="vorticity")
    assert rel_l2(Ew, E) < 1e-12


def test_energy_spectrum_single_mode(vp, O):
    """One planted mode k = (3, -2, 1) (and one at kx = 0): energy in bin round(|k|/Δk) only."""
    u = np.zeros((3, 32, 32, 17), dtype=complex)
    u[:, 1, -2 % 32, 3] = [0.5 + 0.25j, -0.1j, 0.3]
    E = O.energy_spectrum(u, (TWO_PI,) * 3, 20, 1.0)
    n = int(round(math.sqrt(14)))
    assert np.count_nonzero(E) == 1 and E[n] == pytest.approx(np.sum(np.abs(u) ** 2))
    u0 = np.zeros_like(u)
    u0[:, 2, 1, 0] = [1.0, 0, 0]
    u0[:, -2 % 32, -1 % 32, 0] = [1.0, 0, 0]  # Hermitian partner
    E0 = O.energy_spectrum(u0, (TWO_PI,) * 3, 20, 1.0)
    assert E0[2] == pytest.approx(1.0) and np.count_nonzero(E0) == 1  # ½·(1 + 1)


def test_planted_mistake_spectrum_factor(vp, O, monkeypatch):
    """Dropping the Hermitian factor on the kx = 0 plane breaks the Parseval check."""
    fs, p, pd = _ring_cfg1(vp)
    u = O.longrange_fourier(pd, p, "velocity")
    orig = np.where
    calls = {"n": 0}

    def where_no_half(cond, a, b):
        calls["n"] += 1
        if isinstance(a, float) and a == 0.5 and b == 1.0:
            return orig(cond, 1.0, 1.0)
        return orig(cond, a, b)

    monkeypatch.setattr(O.np, "where", where_no_half)
    E = O.energy_spectrum(u, p.Ls, 60, 1.0)
    monkeypatch.undo()
    g = np.fft.irfftn(u, s=(64, 64, 64), axes=(1, 2, 3), norm="forward")
    assert abs(E.sum() - 0.5 * np.mean(np.sum(g ** 2, axis=0))) > 1e-3 * E.sum()


def test_band_field(vp, O):
    """Band-filtered field at the nodes: NUFFT against the direct sum, and a planted mode against 2·Re(c e^{ik·x})."""
    fs, p, pd = _ring_cfg1(vp)
    u = O.longrange_fourier(pd, p, "velocity")
    full = O.band_field(u, p, pd.nodes, 0.0, 100.0, "exact")
    for kmin, kmax in ((2.0, 6.0), (0.0, 9.5), (20.0, 31.0)):
        be = O.band_field(u, p, pd.nodes, kmin, kmax, "exact")
        bn = O.band_field(u, p, pd.nodes, kmin, kmax, "nufft")
        # the type-2 NUFFT error is set by the kernel, relative to the whole field (~3e-7 at m = 4, σ = 1.5)
        assert np.linalg.norm(bn - be) < 1e-6 * np.linalg.norm(full), (kmin, kmax)
    c = np.array([0.3 - 0.2j, 0.1j, -0.05])
    u1 = np.zeros_like(u)
    u1[:, 1, 62, 3] = c  # k = (3, -2, 1)
    x = pd.nodes
    ph = 3 * x[0] - 2 * x[1] + 1 * x[2]
    want = 2 * np.real(c[:, None] * np.exp(1j * ph)[None])
    assert rel_l2(O.band_field(u1, p, x, 3.0, 4.0, "exact"), want) < 1e-13
    assert np.all(O.band_field(u1, p, x, 4.0, 9.0, "exact") == 0)  # |k| = √14 < 4 is outside


# ----------------------------------------------------------------------------
# Background vorticity (unclosed periodic lines with a net end-to-end offset)


def background_lines(vp, Np=64, L=TWO_PI):
    """4×4 array of aligned periodic lines along z with identical helical perturbations
    r(t) = e^{4πit}/100 (test/background_vorticity.jl:38-56): a net mean vorticity."""
    xs = [L * (2 * i + 1) / 8 for i in range(4)]  # range(0, L; length = 9)[2:2:end]
    fs = []
    for x in xs:
        for y in xs:
            S = vp.define_curve(vp.PeriodicLine(r=lambda t: np.exp(4j * np.pi * np.asarray(t)) / 100), scale=L,
                                translate=(x, y, L / 2))
            fs.append(vp.filaments.init(S, Np))
    return fs


def background_params(vp, Ngrid):
    """GaussianSplitting(rtol = 1e-6, Ns) with Γ = 2, a = 1e-8, m = 4 (background_vorticity.jl:18-25,62-64);
    doubling Ns doubles α."""
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, rtol=1e-6, Ns=Ngrid)
    return vp.ParamsBiotSavart.build(np.float64, Γ=2.0, a=1e-8, splitting=spl,
                                     backend_long=vp.HIPLongRangeBackend(σ=1.5, m=4))


def background_alpha_diffs(vp, O, fs):
    """Per-filament max over filaments of |a - b| / max(|a|, |b|) between the Ns = 32 and
    Ns = 64 splittings, for v and ψ (Julia's `a ≈ b` on each filament's vector of nodes)."""
    out = []
    for Ngrid in (32, 64):
        p = background_params(vp, Ngrid)
        pd = vp.add_point_charges(fs, p)
        r = O.compute(pd, p, ("velocity", "streamfunction"))
        out.append((pd, r))
    b = out[0][0].batch
    diffs = {}
    for k in ("velocity", "streamfunction"):
        a1, a2 = out[0][1][k], out[1][1][k]
        worst = 0.0
        for o, n in zip(b.offs, b.Ns):
            x, y = a1[:, o : o + n], a2[:, o : o + n]
            worst = max(worst, np.linalg.norm(x - y) / max(np.linalg.norm(x), np.linalg.norm(y)))
        diffs[k] = worst
    return diffs


def test_background_vorticity_alpha_independence(vp, O):
    """test/background_vorticity.jl:58-80: v within rtol 2e-6 and ψ within 2e-7 between α and 2α."""
    fs = background_lines(vp)
    d = background_alpha_diffs(vp, O, fs)
    print(f"  α-independence with a mean vorticity: v {d['velocity']:.2e}, ψ {d['streamfunction']:.2e}")
    assert d["velocity"] < 2e-6
    assert d["streamfunction"] < 2e-7
    # the correction is what makes ψ independent of α here: it must be large
    p = background_params(vp, 32)
    pd = vp.add_point_charges(fs, p)
    shift = O.background_vorticity_psi(pd, p)
    assert abs(shift[2]) > 1e-3 and abs(shift[0]) < 1e-12 and abs(shift[1]) < 1e-12
    # ⟨ω⟩ = 16 Γ L ẑ / L³ and the shift is −⟨ω⟩/(4α²) (gaussian.jl:241)
    want = -16 * 2.0 * TWO_PI / TWO_PI ** 3 / (4 * p.α ** 2)
    assert abs(shift[2] - want) < 1e-12 * abs(want)


@pytest.mark.parametrize("mistake", ["sign", "double", "missing"])
def test_planted_mistake_background_vorticity(vp, O, monkeypatch, mistake):
    """A flipped sign, a doubled factor or a missing correction breaks the α-independence of ψ."""
    orig = O.background_vorticity_psi
    f = {"sign": -1.0, "double": 2.0, "missing": 0.0}[mistake]
    monkeypatch.setattr(O, "background_vorticity_psi", lambda pd, p: f * orig(pd, p))
    d = background_alpha_diffs(vp, O, background_lines(vp))
    assert d["streamfunction"] > 1e-4, d
    assert d["velocity"] < 2e-6  # the velocity does not see it
