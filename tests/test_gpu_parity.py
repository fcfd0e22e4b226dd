"""GPU parity: libvpbs (through its C ABI) against the FP64 CPU oracle.

Bar (from the reference's own CPU-vs-device parity test, test/ring_collision.jl:153-156):
plain FP64 path ≤ 1e-12 relative L2 against the oracle on identical point data,
and bit-identical short-range neighbour sets (per-node accepted-pair counts and
order-independent hashes). The FP32-bulk path is checked against the FP64 path
with the tolerances stated below.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TWO_PI = 2 * math.pi
Q = ("velocity", "streamfunction")
TOL64 = 1e-12


def _params(vp, T=np.float64, Ns=64, beta=3.5, Ls=TWO_PI, **kws):
    spl = vp.GaussianSplitting.create(Ls=Ls, β=beta, Ns=Ns)
    return vp.ParamsBiotSavart.build(T, Γ=1.0, a=1e-6, Δ=0.25, splitting=spl, **kws)


def _gpu(vp, params, fs, fields=Q, pointdata="device", **kws):
    cache = vp.init_cache(params, pointdata=pointdata)
    out = {k: [np.zeros((len(f), 3)) for f in fs] for k in fields}
    vp.compute_on_nodes_(out, cache, fs, **kws)
    lin = {k: np.concatenate(v, axis=0).T for k, v in out.items()}
    return cache, lin


def _oracle_kw(kws):
    m = {"LIA": kws.get("LIA", True), "shortrange": kws.get("shortrange", True),
         "longrange_on": kws.get("longrange", True)}
    if "integration_cutoff" in kws:
        m["integration_cutoff"] = kws["integration_cutoff"]
    return m


def _pd(vp, cache, fs):
    """Host point data of fs: the oracle's input (built on the host in both point-data modes)."""
    return cache.pointdata if cache.pointdata is not None else vp.add_point_charges(fs, cache.params)


def _check(vp, O, params, fs, fields=Q, tol=TOL64, **kws):
    cache, got = _gpu(vp, params, fs, fields, **kws)
    want = O.compute(_pd(vp, cache, fs), params, fields, **_oracle_kw(kws))
    for k in fields:
        want_k = want[k].copy()
        if k == "streamfunction" and kws.get("shortrange", True) and kws.get("LIA", True) != "only":
            pass  # background correction applied by both (zero for closed filaments)
        err = rel_l2(got[k], want_k)
        print(f"  {k}: rel L2 vs oracle = {err:.3e}", flush=True)
        assert err <= tol, (k, err)
    cache.close()
    return got


@pytest.fixture(scope="module")
def cfg1(vp):
    return vp.tangles.config_tangle(1)


@pytest.fixture(scope="module")
def cfg2(vp):
    return vp.tangles.config_tangle(2)


def test_cfg1_ring_fp64(vp, O, cfg1):
    """One ring, 64 nodes, 64³ grid, β = 3.5 (configuration 1)."""
    _check(vp, O, _params(vp), cfg1)


@pytest.mark.parametrize("variant", [
    dict(LIA="only"), dict(LIA=False), dict(longrange=False), dict(shortrange=False),
    dict(integration_cutoff=1e-3),
])
def test_cfg1_variants(vp, O, cfg1, variant):
    _check(vp, O, _params(vp), cfg1, **variant)


def test_cfg1_no_avoid_explicit_erf(vp, O, cfg1):
    _check(vp, O, _params(vp, avoid_explicit_erf=False), cfg1)


def test_cfg1_truncate_spherical(vp, O, cfg1):
    _check(vp, O, _params(vp, longrange_truncate_spherical=True), cfg1)


def test_cfg2_random_rings_fp64(vp, O, cfg2):
    """64 random rings (≈16k nodes), 128³ grid (configuration 2): values and neighbour sets."""
    params = _params(vp, Ns=128)
    cache, got = _gpu(vp, params, cfg2, pointdata="host", neighbour_stats=True)
    pd = cache.pointdata
    print(f"  Np = {pd.Np}", flush=True)
    assert 14000 < pd.Np < 19000
    want = O.compute(pd, params, Q, neighbour_stats=True)
    for k in Q:
        err = rel_l2(got[k], want[k])
        print(f"  {k}: rel L2 vs oracle = {err:.3e}", flush=True)
        assert err <= TOL64
    counts, hashes = cache.ctx.neighbour_stats()
    oc, oh = want["stats"]
    assert np.array_equal(counts, oc) and np.array_equal(hashes, oh)
    bc, bh = O.neighbour_pairs_bruteforce(pd, params)
    assert np.array_equal(counts, bc) and np.array_equal(hashes, bh)
    print(f"  accepted pairs: {counts.sum()} ({counts.mean():.1f} per node)", flush=True)
    cache.close()


def test_trefoil_noncubic_box(vp, O):
    """Non-cubic periods and odd grid sizes (test/trefoil.jl:393-396)."""
    R = math.pi / 3
    S = vp.define_curve(vp.TrefoilKnot(), scale=R, translate=R)
    fs = [vp.ClosedFilament(S(np.arange(30) / 30), np.zeros(3))]
    Ls = (1.5 * math.pi, 1.5 * math.pi, 2 * math.pi)
    spl = vp.GaussianSplitting.create(Ls=Ls, α=3.0, rcut=1.0, Ns=(45, 45, 61))
    params = vp.ParamsBiotSavart.build(np.float64, Γ=7.4, a=1e-5, splitting=spl)
    _check(vp, O, params, fs)


def test_unclosed_lines_background_vorticity(vp, O):
    """Periodic lines with a net end-to-end offset: ψ gets the background correction (BiotSavart.jl:255-286)."""
    fs = vp.tangles.kelvin_lines(2, TWO_PI / 48, seed=5)
    fs = [fs[0]]  # one line along x: non-zero mean vorticity
    params = _params(vp, Ns=48)
    cache, got = _gpu(vp, params, fs)
    pd = _pd(vp, cache, fs)
    want = O.compute(pd, params, Q)
    for k in Q:
        assert rel_l2(got[k], want[k]) <= TOL64
    assert np.abs(O.background_vorticity_psi(pd, params)).max() > 0
    cache.close()


def test_fp32_bulk_path(vp, O, cfg2):
    """FP32 bulk arithmetic against the FP64 oracle.

    Tolerances: 2e-6 relative L2 for velocity and streamfunction (the NUFFT itself
    is accurate to ~4e-7 at m = 4, σ = 1.5; FP32 rounding of the FFT and of the
    compensated sums stays below that).
    """
    params = _params(vp, T=np.float32, Ns=128)
    cache, got = _gpu(vp, params, cfg2, pointdata="host", neighbour_stats=True)
    p64 = _params(vp, Ns=128)
    want = O.compute(cache.pointdata, p64, Q, neighbour_stats=True)
    for k in Q:
        err = rel_l2(got[k], want[k])
        print(f"  fp32 {k}: rel L2 vs FP64 oracle = {err:.3e}", flush=True)
        assert err < 2e-6, (k, err)
    # the FP32 path tests r² ≤ rcut² in FP64: same pair set as the oracle's
    # local-segment-excluding pass (avoid_explicit_erf = false semantics)
    p64b = _params(vp, Ns=128, avoid_explicit_erf=False)
    ref = O.compute(cache.pointdata, p64b, ("velocity",), longrange_on=False, neighbour_stats=True)["stats"]
    counts, hashes = cache.ctx.neighbour_stats()
    assert np.array_equal(counts, ref[0]) and np.array_equal(hashes, ref[1])
    cache.close()


def test_deterministic_repeat(vp, cfg2):
    """Two evaluations of the same input are bit-identical (deterministic sorts, no float atomics)."""
    params = _params(vp, Ns=128)
    cache = vp.init_cache(params)
    a = vp.velocity_on_nodes(cache, cfg2)
    b = vp.velocity_on_nodes(cache, cfg2)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    cache.close()


def test_resize_between_calls(vp, O, cfg1, cfg2):
    """One context, point sets of different sizes in a row."""
    params = _params(vp, Ns=128)
    cache = vp.init_cache(params)
    for fs in (cfg2, cfg1, cfg2[:5]):
        out = {k: [np.zeros((len(f), 3)) for f in fs] for k in Q}
        vp.compute_on_nodes_(out, cache, fs)
        want = O.compute(_pd(vp, cache, fs), params, Q)
        for k in Q:
            assert rel_l2(np.concatenate(out[k]).T, want[k]) <= TOL64
    cache.close()


def test_error_paths(vp, cfg1):
    params = _params(vp)
    ctx = vp._lib.Context(params)
    with pytest.raises(vp._lib.VpbsError, match="before vpbs_set_points"):
        ctx.compute(vp._lib.VELOCITY)
    with pytest.raises(vp._lib.VpbsError, match="VPBS_VELOCITY"):
        ctx.set_points(vp.add_point_charges(cfg1, params))
        ctx.compute(0)
    ctx.close()
    bad = vp._lib.make_params(params)
    bad.rcut = 3.0  # above M/(2M+1)·L
    import ctypes

    h = ctypes.c_void_p()
    L = vp._lib.load()
    assert L.vpbs_create(ctypes.byref(h), ctypes.byref(bad)) == -1
    assert b"larger than that allowed" in L.vpbs_last_error(None)
    bad = vp._lib.make_params(params)
    bad.abi_version = 99
    assert L.vpbs_create(ctypes.byref(h), ctypes.byref(bad)) == -1
    with pytest.raises(ValueError):
        cache = vp.init_cache(params)
        vp.compute_on_nodes_({"velocity": [np.zeros((3, 3))]}, cache, cfg1)


@pytest.mark.parametrize("sigma,m", [(1.25, 5), (2.0, 3), (1.5, 6)])
def test_nufft_kernel_widths(vp, O, cfg1, sigma, m):
    """Other oversampling factors / half-supports (generic spreading and interpolation paths)."""
    params = _params(vp, backend_long=vp.HIPLongRangeBackend(σ=sigma, m=m))
    _check(vp, O, params, cfg1)


def test_tiny_grid_few_tiles(vp, O):
    """Ns = 10 → the smallest NUFFT grid (32³): one spreading tile along x, four along y and z."""
    fs = vp.tangles.random_rings(3, 0.8, 1.2, TWO_PI / 24, seed=2)
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, α=1.0, rcut=2.0, Ns=10)
    params = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl)
    _check(vp, O, params, fs)


def test_lia_segment_fraction(vp, O, cfg2):
    """lia_segment_fraction = 0.2 with GL(8) near the singularity (host-side local integrals on both sides)."""
    params = _params(vp, Ns=128, lia_segment_fraction=0.2, quadrature_near_singularity=vp.GaussLegendre(8))
    _check(vp, O, params, cfg2)


# --------------------------------------------------------------------------- #
# Device-side point data (vpbs_set_filaments) against the host add_point_charges


def _pointdata_cases(vp):
    R = math.pi / 3
    tref = vp.define_curve(vp.TrefoilKnot(), scale=R, translate=R)
    five = [vp.filaments.init(vp.define_curve(vp.Ring(), scale=0.4, translate=(1.0, 2.0, 3.0)), 5)]
    far = [vp.ClosedFilament(f.nodes + np.array([3 * TWO_PI, -2 * TWO_PI, 0.0]), f.offset)
           for f in vp.tangles.random_rings(4, 0.5, 1.0, TWO_PI / 40, seed=3)]
    return {
        "ring_gl3": (vp.tangles.single_ring(), {}),
        "random_rings_gl3": (vp.tangles.config_tangle(2), {}),
        "trefoil_gl2_noncubic": ([vp.filaments.init(tref, 30)], dict(quadrature=vp.GaussLegendre(2),
                                  Ls=(1.5 * math.pi, 1.5 * math.pi, 2 * math.pi))),
        "kelvin_lines_gl4": (vp.tangles.kelvin_lines(3, TWO_PI / 48, seed=5), dict(quadrature=vp.GaussLegendre(4))),
        "five_nodes_gl3": (five, {}),
        "straight_segments": (vp.tangles.random_rings(6, 0.5, 1.2, TWO_PI / 32, seed=9),
                              dict(quadrature=vp.NoQuadrature())),
        "straight_segments_lia_fraction": (vp.tangles.kelvin_lines(2, TWO_PI / 40, seed=1),
                                           dict(quadrature=vp.NoQuadrature(), lia_segment_fraction=0.4)),
        "lia_fraction_gl3": (vp.tangles.config_tangle(2)[:8], dict(lia_segment_fraction=0.2)),
        "unfolded_positions": (far, {}),
    }


def _pd_params(vp, fs, Ls=TWO_PI, **kws):
    spl = vp.GaussianSplitting.create(Ls=Ls, β=3.5, Ns=48)
    return vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl, **kws)


@pytest.mark.parametrize("case", ["ring_gl3", "random_rings_gl3", "trefoil_gl2_noncubic", "kelvin_lines_gl4",
                                  "five_nodes_gl3", "straight_segments", "straight_segments_lia_fraction",
                                  "lia_fraction_gl3", "unfolded_positions"])
def test_device_pointdata_matches_host(vp, case):
    """Point data built on the GPU equals add_point_charges on the host.

    Bar: bit-identical. The device evaluates the same IEEE operations in the
    same order as the host restatement, without FMA contraction, and FP64
    division and square root are correctly rounded on both sides. (A 1-ulp
    difference in a point lying on a NUFFT grid plane changes the long-range
    velocity by up to e^{-β} relative, so identical inputs matter for the
    parity of everything downstream.) Invariant: for each filament the
    charges sum to the end-to-end offset, ∑ w·Δt·S′ = S(T) - S(0) (GL with
    ≥ 3 points integrates the quartic S′ exactly; straight segments telescope).
    """
    fs, kw = _pointdata_cases(vp)[case]
    Ls = kw.pop("Ls", TWO_PI)
    params = _pd_params(vp, fs, Ls=Ls, **kw)
    print(f"  case {case}: {len(fs)} filaments, Np = {sum(len(f) for f in fs)}", flush=True)
    host = vp.add_point_charges(fs, params)
    ctx = vp._lib.Context(params)
    ctx.set_filaments(vp.FilamentBatch(fs), params)
    dev = ctx.get_points()
    ctx.close()
    assert np.array_equal(dev["node_idx_prev"], host.node_idx_prev)
    assert np.array_equal(dev["nodes"], host.nodes)
    for name in ("sp", "spp", "dlm", "dlp", "points", "charges"):
        got, want = dev[name], getattr(host, name)
        err = np.abs(got - want).max() / np.abs(want).max()
        print(f"    {name}: max |dev - host| / scale = {err:.2e}", flush=True)
        assert np.array_equal(got, want), (name, err)
    b = host.batch
    nq = host.nquad
    if nq >= 3 or isinstance(params.quadrature, vp.NoQuadrature):
        for f in range(b.Nf):
            a, n = b.offs[f], b.Ns[f]
            tot = dev["charges"][:, nq * a: nq * (a + n)].sum(axis=1)
            assert np.allclose(tot, b.Xoff[f], rtol=0, atol=1e-12 * max(params.Ls)), (f, tot, b.Xoff[f])


def test_device_pointdata_compute(vp, O, cfg2):
    """Full evaluation with device-built point data equals the host-built one (FP64, 1e-12)."""
    params = _params(vp, Ns=128)
    _, got_dev = _gpu(vp, params, cfg2, pointdata="device")
    _, got_host = _gpu(vp, params, cfg2, pointdata="host")
    for k in Q:
        err = rel_l2(got_dev[k], got_host[k])
        print(f"  {k}: rel L2 device vs host point data = {err:.3e}", flush=True)
        assert err <= TOL64


def test_device_pointdata_errors(vp, cfg1):
    """Argument checks of vpbs_set_filaments: every one raises before any kernel runs, except the
    non-finite-position case, which the folded-range check after the build reports."""
    params = _params(vp)
    ctx = vp._lib.Context(params)
    L = vp._lib.load()
    b = vp.FilamentBatch(cfg1)
    X = np.ascontiguousarray(b.X.T)
    c = np.ascontiguousarray(b.c.T)
    xoff = np.ascontiguousarray(b.Xoff.T)
    zs, ws = (np.ascontiguousarray(a) for a in params.quadrature.rule())
    P = vp._lib._ptr

    def call(counts, X=X, T=b.T, zs=zs, ws=ws, g=math.nan):
        counts = np.asarray(counts, dtype=np.int64)
        return L.vpbs_set_filaments(ctx._h, len(counts), P(counts), P(X), P(b.ts), P(np.asarray(T)), P(xoff),
                                    P(c), P(zs), P(ws), g)

    def msg():
        return L.vpbs_last_error(ctx._h).decode()

    assert call([4]) == -1 and "at least 5" in msg()
    assert call(b.Ns, T=np.array([0.0])) == -1 and "period" in msg()
    assert call(b.Ns, zs=None) == -1 and "both" in msg()
    assert call(b.Ns, g=1.5) == -1 and "lia_segment_fraction" in msg()
    assert L.vpbs_set_filaments(ctx._h, 0, None, None, None, None, None, None, None, None, math.nan) == -1
    Xbad = X.copy()
    Xbad[1, 3] = np.nan
    assert call(b.Ns, X=Xbad) == -1 and "not folded" in msg()
    with pytest.raises(vp._lib.VpbsError, match="before vpbs_set_points"):
        ctx.compute(vp._lib.VELOCITY)
    ctx.set_filaments(b, params)  # a good call after the failed ones
    v, _ = ctx.compute(vp._lib.VELOCITY)
    assert v.shape == (3, b.Np) and np.isfinite(v).all()
    ctx.close()
    p1 = _params(vp, quadrature=vp.GaussLegendre(2))
    ctx = vp._lib.Context(p1)
    with pytest.raises(vp._lib.VpbsError, match="nquad = 1"):
        ctx.set_filaments(b, _params(vp, quadrature=vp.NoQuadrature()))
    ctx.close()
