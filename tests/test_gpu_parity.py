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


def test_background_vorticity_alpha_independence(vp, O):
    """The product's background-vorticity correction (host side of compute_on_nodes_) pinned by the
    α-independence of test/background_vorticity.jl:58-80: a 4×4 array of unclosed lines, Ns = 32
    against Ns = 64 (α doubles); per filament v within rtol 2e-6 and ψ within 2e-7; each run also
    ≤ 1e-12 against the oracle."""
    from test_oracle import background_lines, background_params

    fs = background_lines(vp)
    runs = []
    for Ngrid in (32, 64):
        p = background_params(vp, Ngrid)
        cache, got = _gpu(vp, p, fs)
        want = O.compute(_pd(vp, cache, fs), p, Q)
        for k in Q:
            assert rel_l2(got[k], want[k]) <= TOL64, (Ngrid, k)
        runs.append(got)
        cache.close()
    b = vp.FilamentBatch(fs)
    for k, tol in (("velocity", 2e-6), ("streamfunction", 2e-7)):
        worst = 0.0
        for o, n in zip(b.offs, b.Ns):
            x, y = runs[0][k][:, o : o + n], runs[1][k][:, o : o + n]
            worst = max(worst, np.linalg.norm(x - y) / max(np.linalg.norm(x), np.linalg.norm(y)))
        print(f"  {k}: α-independence {worst:.2e} (tolerance {tol})", flush=True)
        assert worst < tol


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


def test_pair_candidate_count(vp, cfg2):
    """vpbs_timings.pair_candidates after a VPBS_NEIGHBOUR_STATS call lies between the accepted
    pair count and the number of sources in the (2M+1)³ neighbour cells of every node (periodic
    3D histogram of the sources, cell index trunc(x/L·N) as in the cell-list build): the walk
    visits only the cells of that cube that the rcut sphere around the node reaches (about
    2/3 of them for M = 2); 0 without the flag."""
    params = _params(vp, Ns=128)
    pd = vp.add_point_charges(cfg2, params)
    ctx = vp._lib.Context(params)
    ctx.set_points(pd)
    ctx.compute(vp._lib.VELOCITY | vp._lib.NEIGHBOUR_STATS, fetch=False)
    got = ctx.timings()["pair_candidates"]
    counts, _ = ctx.neighbour_stats()
    _, Nc = ctx.sizes()
    ctx.compute(vp._lib.VELOCITY, fetch=False)
    assert ctx.timings()["pair_candidates"] == 0
    ctx.close()
    M = params.backend_short.nsubdiv
    L = np.asarray(params.Ls, dtype=np.float64)
    cell = lambda X: [((X[d] / L[d]) * Nc[d]).astype(np.int64) for d in range(3)]  # noqa: E731 (X: [3][n])
    sx, sy, sz = cell(np.asarray(pd.points))
    H = np.zeros(Nc[::-1], dtype=np.int64)
    np.add.at(H, (sz, sy, sx), 1)
    S = np.zeros_like(H)
    for dz in range(-M, M + 1):
        for dy in range(-M, M + 1):
            for dx in range(-M, M + 1):
                S += np.roll(H, (dz, dy, dx), axis=(0, 1, 2))
    nx, ny, nz = cell(np.asarray(pd.nodes))
    want = int(S[nz, ny, nx].sum())
    print(f"  candidates {got} (cube {want}, {got / want:.2f}), accepted {int(counts.sum())}", flush=True)
    assert want > got >= counts.sum() > 0 and got < 0.8 * want


@pytest.mark.parametrize("M,Ns,cells", [(1, 26, 3), (1, 34, 4), (2, 22, 5), (2, 26, 6), (3, 24, 8), (4, 20, 9),
                                         (4, 22, 10), (4, 26, 12)])
def test_few_cells_min_image(vp, O, cfg2, M, Ns, cells):
    """Neighbour sets and values with 2M+1 and 2M+2 cells per direction. The pair kernel skips
    the minimum-image fold for nodes whose neighbour cells do not wrap when there are at
    least 2M+2 cells; with 2M+1 cells the neighbour range covers the whole box and every node
    folds. Both must give the brute-force pair sets and the oracle's values. M = 4 has 81
    neighbour cell rows, more than the 64 lanes of the pair kernel's wavefront."""
    fs = cfg2[:16]
    params = _params(vp, Ns=Ns, backend_short=vp.HIPShortRangeBackend(nsubdiv=M))
    assert math.floor(TWO_PI * M / params.rcut) == cells
    cache, got = _gpu(vp, params, fs, ("velocity",), pointdata="host", neighbour_stats=True)
    pd = cache.pointdata
    counts, hashes = cache.ctx.neighbour_stats()
    bc, bh = O.neighbour_pairs_bruteforce(pd, params)
    print(f"  M = {M}, {cells} cells, Np = {pd.Np}, accepted {int(counts.sum())}", flush=True)
    assert np.array_equal(counts, bc) and np.array_equal(hashes, bh)
    want = O.compute(pd, params, ("velocity",))
    err = rel_l2(got["velocity"], want["velocity"])
    print(f"  velocity: rel L2 vs oracle = {err:.3e}", flush=True)
    assert err <= TOL64
    cache.close()


@pytest.mark.parametrize("M,Ns", [(2, 128), (3, 128), (2, 24)], ids=["M2", "M3", "M2_few_cells"])
def test_tiled_pair_kernel(vp, O, cfg2, monkeypatch, M, Ns):
    """The tiled pair kernel (VPBS_PAIR=tile at context creation): brute-force pair sets and the
    oracle's values, FP64; with fewer than 2 + 2M cells per direction the library falls back to
    one wavefront per target."""
    fs = cfg2[:16]
    monkeypatch.setenv("VPBS_PAIR", "tile")
    params = _params(vp, Ns=Ns, backend_short=vp.HIPShortRangeBackend(nsubdiv=M))
    cache, got = _gpu(vp, params, fs, Q, pointdata="host", neighbour_stats=True)
    monkeypatch.delenv("VPBS_PAIR")
    pd = cache.pointdata
    counts, hashes = cache.ctx.neighbour_stats()
    bc, bh = O.neighbour_pairs_bruteforce(pd, params)
    assert np.array_equal(counts, bc) and np.array_equal(hashes, bh)
    want = O.compute(pd, params, Q)
    for k in Q:
        err = rel_l2(got[k], want[k])
        print(f"  M = {M}, Ns = {Ns}, cells {cache.ctx.sizes()[1]}: {k} rel L2 vs oracle = {err:.3e}", flush=True)
        assert err <= TOL64
    cache.close()


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_deterministic_repeat(vp, cfg2, T):
    """Two evaluations of the same input are bit-identical (deterministic sorts, no float atomics;
    the slab interpolation sums each node's slab parts in a fixed order)."""
    params = _params(vp, T=T, Ns=128)
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
    """lia_segment_fraction = 0.2 with GL(8) near the singularity: local integrals on the device (device point
    data) against the oracle's host restatement."""
    params = _params(vp, Ns=128, lia_segment_fraction=0.2, quadrature_near_singularity=vp.GaussLegendre(8))
    _check(vp, O, params, cfg2)


@pytest.mark.parametrize("variant", [dict(), dict(LIA=False), dict(LIA="only"), dict(longrange=False)])
def test_local_integrals_device_vs_host(vp, cfg2, variant):
    """Device local integrals (device point data) equal the host ones (host point data): FP64 1e-13,
    for every LIA setting (BiotSavart.jl:570-576 adds them unless LIA = Val(:only))."""
    params = _params(vp, Ns=128, lia_segment_fraction=0.3, quadrature_near_singularity=vp.GaussLegendre(5))
    _, got_dev = _gpu(vp, params, cfg2, pointdata="device", **variant)
    _, got_host = _gpu(vp, params, cfg2, pointdata="host", **variant)
    for k in Q:
        err = rel_l2(got_dev[k], got_host[k])
        print(f"  {variant} {k}: device vs host local integrals {err:.2e}", flush=True)
        assert err < 1e-13


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
                                    P(c), P(zs), P(ws), g, 0, None, None)

    def msg():
        return L.vpbs_last_error(ctx._h).decode()

    assert call([4]) == -1 and "at least 5" in msg()
    assert call(b.Ns, T=np.array([0.0])) == -1 and "period" in msg()
    assert call(b.Ns, zs=None) == -1 and "both" in msg()
    assert call(b.Ns, g=1.5) == -1 and "lia_segment_fraction" in msg()
    assert L.vpbs_set_filaments(ctx._h, 0, None, None, None, None, None, None, None, None, math.nan, 0, None,
                                None) == -1
    assert L.vpbs_set_filaments(ctx._h, 1, P(np.asarray(b.Ns, dtype=np.int64)), P(X), P(b.ts), P(b.T), P(xoff),
                                P(c), P(zs), P(ws), 0.3, 17, P(zs), P(ws)) == -1 and "near-singularity" in msg()
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


# --------------------------------------------------------------------------- #
# Long-range field in Fourier space and its device-side consumers


def _uhat_case(vp, fs, Ns=64, T=np.float64, **kws):
    params = _params(vp, T=T, Ns=Ns, **kws)
    cache = vp.init_cache(params)  # keep_fourier=True by default
    vs = [np.zeros((len(f), 3)) for f in fs]
    vp.compute_on_nodes_({"velocity": vs}, cache, fs)
    return params, cache, vs


@pytest.mark.parametrize("variant", ["cfg1", "cfg2", "truncate_spherical", "trefoil_noncubic"])
def test_uhat_matches_oracle(vp, O, cfg1, cfg2, variant):
    """Kept unsmoothed velocity v̂ = i k × ω̂/k² on the Fourier grid vs the oracle (FP64, 1e-12)."""
    if variant == "trefoil_noncubic":
        R = math.pi / 3
        S = vp.define_curve(vp.TrefoilKnot(), scale=R, translate=R)
        fs = [vp.ClosedFilament(S(np.arange(30) / 30), np.zeros(3))]
        spl = vp.GaussianSplitting.create(Ls=(1.5 * math.pi, 1.5 * math.pi, 2 * math.pi), α=3.0, rcut=1.0,
                                         Ns=(45, 45, 61))
        params = vp.ParamsBiotSavart.build(np.float64, Γ=7.4, a=1e-5, splitting=spl)
        cache = vp.init_cache(params)
        vp.compute_on_nodes_({"velocity": [np.zeros((30, 3))]}, cache, fs)
    else:
        fs = cfg2 if variant == "cfg2" else cfg1
        kw = {"longrange_truncate_spherical": True} if variant == "truncate_spherical" else {}
        params, cache, _ = _uhat_case(vp, fs, Ns=128 if variant == "cfg2" else 64, **kw)
    got = vp.get_longrange_field_fourier(cache)
    assert got["state"] == "velocity"
    want = O.longrange_fourier(vp.add_point_charges(fs, params), params, "velocity")
    err = rel_l2(np.stack(got["field"]), want)
    print(f"  {variant}: uhat rel L2 vs oracle = {err:.3e}", flush=True)
    assert err <= TOL64
    cache.close()


def test_uhat_streamfunction_state(vp, O, cfg1):
    """Streamfunction only: the kept field is ψ̂ = ω̂/k² (longrange.jl:102-113)."""
    params = _params(vp)
    cache = vp.init_cache(params)
    vp.compute_on_nodes_({"streamfunction": [np.zeros((len(f), 3)) for f in cfg1]}, cache, cfg1)
    got = vp.get_longrange_field_fourier(cache)
    assert got["state"] == "streamfunction"
    want = O.longrange_fourier(vp.add_point_charges(cfg1, params), params, "streamfunction")
    assert rel_l2(np.stack(got["field"]), want) <= TOL64
    with pytest.raises(RuntimeError, match="not supported"):
        vp.diagnostics.energy_spectrum(cache)
    # an evaluation without the long range leaves no field
    vp.compute_on_nodes_({"velocity": [np.zeros((len(f), 3)) for f in cfg1]}, cache, cfg1, longrange=False)
    assert cache.ctx.uhat_quantity() is None
    with pytest.raises(RuntimeError):
        vp.get_longrange_field_fourier(cache)
    cache.close()


def test_energy_spectrum_device(vp, O, cfg2):
    """Device energy spectrum vs the oracle's binning of the oracle field (FP64, 1e-12); a finer Δk too."""
    params, cache, _ = _uhat_case(vp, cfg2, Ns=128)
    ks, Ek = vp.diagnostics.energy_spectrum(cache)
    assert len(ks) == 65 and ks[0] == 0
    u = O.longrange_fourier(vp.add_point_charges(cfg2, params), params, "velocity")
    want = O.energy_spectrum(u, params.Ls, len(ks), ks[1] - ks[0])
    err = rel_l2(Ek, want)
    print(f"  spectrum rel L2 vs oracle = {err:.3e}; E(k=1..4) = {Ek[1:5]}", flush=True)
    assert err <= TOL64
    Ek2 = cache.ctx.energy_spectrum(300, 0.37)
    assert rel_l2(Ek2, O.energy_spectrum(u, params.Ls, 300, 0.37)) <= TOL64
    cache.close()


def test_band_interpolation(vp, O, cfg2):
    """Band-filtered field at the nodes: vs the oracle's NUFFT restatement (1e-12) and the direct sum (NUFFT tol.)."""
    params, cache, _ = _uhat_case(vp, cfg2, Ns=128)
    pd = vp.add_point_charges(cfg2, params)
    u = O.longrange_fourier(pd, params, "velocity")
    full = O.band_field(u, params, pd.nodes, 0.0, 1e3, "exact")
    for kmin, kmax in ((0.0, 4.0), (2.0, 6.0), (10.0, 30.0)):
        got = cache.ctx.interpolate_band(kmin, kmax)
        err = rel_l2(got, O.band_field(u, params, pd.nodes, kmin, kmax, "nufft"))
        errx = np.linalg.norm(got - O.band_field(u, params, pd.nodes, kmin, kmax, "exact")) / np.linalg.norm(full)
        print(f"  band [{kmin}, {kmax}]: vs oracle NUFFT {err:.2e}, vs direct sum {errx:.2e}", flush=True)
        assert err <= TOL64 and errx < 1e-6
    cache.close()


def test_uhat_fp32(vp, O, cfg2):
    """FP32 path: kept field and spectrum against the FP64 oracle within FP32 tolerance."""
    params, cache, _ = _uhat_case(vp, cfg2, Ns=128, T=np.float32)
    p64 = _params(vp, Ns=128)
    u = O.longrange_fourier(vp.add_point_charges(cfg2, p64), p64, "velocity")
    err = rel_l2(np.stack(vp.get_longrange_field_fourier(cache)["field"]), u)
    ks, Ek = vp.diagnostics.energy_spectrum(cache)
    errE = rel_l2(Ek, O.energy_spectrum(u, p64.Ls, len(ks), ks[1] - ks[0]))
    print(f"  fp32 uhat {err:.2e}, spectrum {errE:.2e}", flush=True)
    assert err < 2e-6 and errE < 2e-6
    cache.close()


def test_forcing_and_dissipation(vp, cfg2):
    """FourierBandForcingBS / SmallScaleDissipationBS on the device band field: invariants of the formulas."""
    params, cache, vs = _uhat_case(vp, cfg2, Ns=128)
    b = vp.FilamentBatch(cfg2)
    st = b.evaluate(b.fid, b.local, np.zeros(b.Np), 1)
    stt = b.evaluate(b.fid, b.local, np.zeros(b.Np), 2)
    sp = st / np.linalg.norm(st, axis=1)[:, None]
    F = vp.forcing
    vf = np.concatenate(F.forcing_velocity(F.FourierBandForcingBS(kmin=1.5, kmax=4.5, α=0.3), cache, cfg2))
    assert np.abs(np.sum(vf * sp, axis=1)).max() < 1e-12 * np.abs(vf).max()  # ⟂ tangent
    # α term: vf = α s′ × (band/2)
    band = cache.ctx.interpolate_band(1.5, 4.5).T
    assert rel_l2(vf, 0.3 * np.cross(sp, 0.5 * band)) < 1e-14
    # ε_target mode: rescaled so that the estimated injection rate is ε_target
    eps = 2e-3
    vt = np.concatenate(F.forcing_velocity(F.FourierBandForcingBS(kmin=1.5, kmax=4.5, ε_target=eps), cache, cfg2))
    assert rel_l2(vt, vf * np.sum(vt * vf) / np.sum(vf * vf)) < 1e-12  # parallel to the α forcing
    # modify_length = false: along the binormal s′ × s″
    vb = np.concatenate(F.forcing_velocity(F.FourierBandForcingBS(kmin=1.5, kmax=4.5, α=1.0, modify_length=False),
                                           cache, cfg2))
    bn = np.cross(sp, np.cross(sp, np.cross(stt, sp)))
    assert np.linalg.norm(np.cross(vb, bn)) < 1e-10 * np.linalg.norm(vb) * np.linalg.norm(bn)
    # dissipation: -s′ × (v - v<), with v< the low-pass velocity
    vd = np.concatenate(F.dissipation_velocity(F.SmallScaleDissipationBS(kdiss=20.0, α=0.5), cache, cfg2, vs))
    vlow = cache.ctx.interpolate_band(0.0, 20.0).T
    assert rel_l2(vd, 0.5 * np.cross(np.concatenate(vs) - vlow, sp)) < 1e-14
    with pytest.raises(ValueError, match="too large"):
        F.dissipation_velocity(F.SmallScaleDissipationBS(kdiss=70.0, α=0.5), cache, cfg2, vs)
    cache.close()


# --------------------------------------------------------------------------- #
# Device-resident time stepping (vpbs_advance)


def _oracle_velocity(vp, O, params):
    def velocity(b):
        pd = vp.add_point_charges(b.fs, params, batch=b)
        return O.compute(pd, params, ("velocity",))["velocity"].T
    return velocity


# Time steps below the Kelvin-wave stability limit at the node spacing δ: the
# fastest mode, k = π/δ, oscillates at ω ≈ Γ k²/(4π) ln(1/(ka)) (≈ 7e3 for the
# 64-node ring, ≈ 1.3e4 for rings with δ = 2π/256); with ω·Δt ≫ 1 rounding
# differences between two equal implementations grow exponentially.


def test_device_stepper_vs_oracle(vp, O, cfg1):
    """Two SSPRK33 steps of the 64-node ring on the device vs the host stepper driven by the FP64 oracle (1e-10)."""
    params = _params(vp)
    dt = 1e-4
    b0 = vp.FilamentBatch(cfg1)
    want = b0
    for _ in range(2):
        want = vp.timestepping.ssprk33_step(want, dt, _oracle_velocity(vp, O, params))
    cache = vp.init_cache(params)
    st = vp.timestepping.DeviceStepper(cache, b0)
    st.advance(dt, 2)
    got = cache.ctx.get_filaments()
    err = np.abs(got["X"] - want.X).max() / np.abs(want.X - b0.X).max()
    print(f"  displacement error rel. to the displacement: {err:.2e}", flush=True)
    assert err < 1e-10
    assert np.allclose(got["knots"], want.ts, rtol=0, atol=1e-13)
    assert np.allclose(got["coefs"], want.c, rtol=0, atol=1e-12)
    cache.close()


@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_device_stepper_vs_host_stepper(vp, cfg2, T):
    """Three steps of 16 rings: device-resident stepping vs host stepping with the same library velocity."""
    fs = cfg2[:16]
    params = _params(vp, T=T, Ns=128)
    dt = 5e-5
    hcache = vp.init_cache(params, pointdata="host", keep_fourier=False)
    b = vp.FilamentBatch(fs)
    for _ in range(3):
        b = vp.timestepping.ssprk33_step(b, dt, vp.timestepping.cache_velocity(hcache))
    dcache = vp.init_cache(params)
    st = vp.timestepping.DeviceStepper(dcache, vp.FilamentBatch(fs))
    st.advance(dt, 3)
    X = st.nodes()
    X0 = np.concatenate([f.nodes for f in fs])
    err = np.abs(X - b.X).max() / np.abs(b.X - X0).max()
    print(f"  {T.__name__}: device vs host stepper, error rel. to displacement {err:.2e}", flush=True)
    assert err < (1e-10 if T is np.float64 else 1e-6)
    # the final point data is that of the final state
    pd = dcache.ctx.get_points()
    assert np.abs(pd["nodes"] - vp.filaments.fold_coordinates_periodic(X.T.ravel(), 2 * math.pi).reshape(3, -1)).max() < 1e-15
    hcache.close()
    dcache.close()


def test_device_stepper_errors(vp, cfg1):
    params = _params(vp)
    ctx = vp._lib.Context(params)
    with pytest.raises(vp._lib.VpbsError, match="vpbs_set_filaments"):
        ctx.advance(0.01, 1)
    ctx.set_points(vp.add_point_charges(cfg1, params))
    with pytest.raises(vp._lib.VpbsError, match="vpbs_set_filaments"):
        ctx.advance(0.01, 1)
    ctx.set_filaments(vp.FilamentBatch(cfg1), params)
    with pytest.raises(vp._lib.VpbsError, match="positive"):
        ctx.advance(-1.0, 1)
    ctx.advance(0.01, 0)  # no steps: state unchanged
    assert np.array_equal(ctx.get_filaments()["X"], vp.FilamentBatch(cfg1).X)
    ctx.close()


# ---------------------------------------------------------------------------
# Transform paths: the fused hand-written FFTs (default, every test above; on 192, 384
# and (FP32) 768-cell x-grids with the inverse x transform inside the interpolation), the same FFTs with the interpolation in its own kernel ("separate"),
# and rocFFT.

@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_fft_paths_agree(vp, O, cfg2, T):
    """Velocity and streamfunction from the fused transforms, the separate-interpolation variant and
    rocFFT on cfg2; FP64 also vs the oracle."""
    outs = {}
    for fft in ("fused", "separate", "rocfft"):
        params = _params(vp, T=T, Ns=128, backend_long=vp.HIPLongRangeBackend(fft=fft))
        cache, outs[fft] = _gpu(vp, params, cfg2, longrange=True, shortrange=False, LIA=False)
        cache.close()
    tol = 1e-12 if T is np.float64 else 2e-6
    for k in Q:
        for fft in ("fused", "separate"):
            err = rel_l2(outs[fft][k], outs["rocfft"][k])
            print(f"  {T.__name__} {k}: {fft} vs rocFFT {err:.2e}", flush=True)
            assert err <= tol
    if T is np.float64:
        params = _params(vp, Ns=128)
        want = O.compute(vp.add_point_charges(cfg2, params), params, Q, LIA=False, shortrange=False)
        for k in Q:
            assert rel_l2(outs["fused"][k], want[k]) <= TOL64


@pytest.mark.parametrize("variant", ["cfg1", "trefoil_noncubic", "tiny"])
def test_rocfft_path(vp, O, cfg1, variant):
    """The rocFFT transform path against the oracle (FP64, 1e-12)."""
    bl = vp.HIPLongRangeBackend(fft="rocfft")
    if variant == "cfg1":
        _check(vp, O, _params(vp, backend_long=bl), cfg1)
    elif variant == "tiny":
        fs = vp.tangles.random_rings(3, 0.8, 1.2, TWO_PI / 24, seed=2)
        spl = vp.GaussianSplitting.create(Ls=TWO_PI, α=1.0, rcut=2.0, Ns=10)
        _check(vp, O, vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=spl, backend_long=bl), fs)
    else:
        R = math.pi / 3
        S = vp.define_curve(vp.TrefoilKnot(), scale=R, translate=R)
        fs = [vp.ClosedFilament(S(np.arange(30) / 30), np.zeros(3))]
        spl = vp.GaussianSplitting.create(Ls=(1.5 * math.pi, 1.5 * math.pi, 2 * math.pi), α=3.0, rcut=1.0,
                                         Ns=(45, 45, 61))
        _check(vp, O, vp.ParamsBiotSavart.build(np.float64, Γ=7.4, a=1e-5, splitting=spl, backend_long=bl), fs)


def test_rocfft_path_uhat_and_band(vp, O, cfg2):
    """rocFFT path: kept field and band interpolation vs the oracle (same layout of the kept field)."""
    params, cache, _ = _uhat_case(vp, cfg2, Ns=128, backend_long=vp.HIPLongRangeBackend(fft="rocfft"))
    pd = vp.add_point_charges(cfg2, params)
    u = O.longrange_fourier(pd, params, "velocity")
    assert rel_l2(np.stack(vp.get_longrange_field_fourier(cache)["field"]), u) <= TOL64
    got = cache.ctx.interpolate_band(2.0, 6.0)
    assert rel_l2(got, O.band_field(u, params, pd.nodes, 2.0, 6.0, "nufft")) <= TOL64
    cache.close()


@pytest.mark.parametrize("Ns", [256, 512])
def test_fft_paths_agree_static_lengths(vp, Ns):
    """Grids 384³ and 768³ (compile-time transform plans): fused and separate-interpolation vs rocFFT
    long range on a few rings (most slabs without nodes), FP64 to 1e-12 and FP32 to 2e-6."""
    fs = vp.tangles.random_rings(8, 0.5, 1.0, TWO_PI / 128, seed=11)
    for T, tol in ((np.float64, 1e-12), (np.float32, 2e-6)):
        outs = {}
        for fft in ("fused", "separate", "rocfft"):
            params = _params(vp, T=T, Ns=Ns, backend_long=vp.HIPLongRangeBackend(fft=fft))
            cache, outs[fft] = _gpu(vp, params, fs, shortrange=False, LIA=False)
            cache.close()
        for k in Q:
            for fft in ("fused", "separate"):
                err = rel_l2(outs[fft][k], outs["rocfft"][k])
                print(f"  Ns={Ns} {T.__name__} {k}: {fft} vs rocFFT {err:.2e}", flush=True)
                assert err <= tol


@pytest.mark.timeout(900)
@pytest.mark.parametrize("T,tol", [(np.float32, 2e-6), (np.float64, 1e-12)], ids=["f32", "f64"])
def test_fft_paths_agree_1536(vp, T, tol):
    """1024³ Fourier grid (1536³ NUFFT grid, the compile-time 1536-point plans of the 4M-node
    configuration): fused transforms with the slab interpolation (FP32) or the separate
    interpolation (FP64) against rocFFT, long-range velocity of a few rings. One context at a time
    (FP64 needs ~185 GB)."""
    fs = vp.tangles.random_rings(6, 0.5, 1.0, TWO_PI / 256, seed=13)
    batch = vp.FilamentBatch(fs)
    L = vp._lib
    outs = {}
    for fft in ("fused", "rocfft"):
        params = _params(vp, T=T, Ns=1024, backend_long=vp.HIPLongRangeBackend(fft=fft))
        ctx = L.Context(params)
        ctx.set_filaments(batch, params)
        outs[fft], _ = ctx.compute(L.VELOCITY | L.NO_SHORTRANGE)
        grid, _ = ctx.sizes()
        ctx.close()
    assert tuple(grid) == (1536, 1536, 1536)
    err = rel_l2(outs["fused"], outs["rocfft"])
    print(f"  1536^3 {T.__name__}: fused vs rocFFT {err:.2e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("Ns", [127, (128, 120, 100)], ids=["odd_nkx", "noncubic"])
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_slab_interpolation_shapes(vp, O, T, Ns):
    """The slab interpolation (inverse x transform fused with the interpolation) on a 192-cell
    x-grid with an odd number of retained kx (63: single-value loads) and on a 192×192×160 grid
    (non-cubic slab counts): fused vs the separate interpolation and rocFFT (FP64 1e-12, FP32
    2e-6); FP64 also against the oracle."""
    fs = vp.tangles.random_rings(8, 0.5, 1.0, TWO_PI / 128, seed=5)
    outs = {}
    for fft in ("fused", "separate", "rocfft"):
        params = _params(vp, T=T, Ns=Ns, backend_long=vp.HIPLongRangeBackend(fft=fft))
        cache, outs[fft] = _gpu(vp, params, fs, shortrange=False, LIA=False)
        cache.close()
    tol = 1e-12 if T is np.float64 else 2e-6
    for k in Q:
        for fft in ("fused", "separate"):
            err = rel_l2(outs[fft][k], outs["rocfft"][k])
            print(f"  Ns={Ns} {T.__name__} {k}: {fft} vs rocFFT {err:.2e}", flush=True)
            assert err <= tol
    if T is np.float64:
        params = _params(vp, Ns=Ns)
        want = O.compute(vp.add_point_charges(fs, params), params, Q, LIA=False, shortrange=False)
        for k in Q:
            assert rel_l2(outs["fused"][k], want[k]) <= TOL64


def test_empty_and_minimal_inputs(vp, O):
    """No filaments returns the empty outputs; a five-node filament (the fewest a closed quintic spline
    takes) and a mix of very different filament sizes match the oracle (FP64, 1e-12)."""
    params = _params(vp)
    cache = vp.init_cache(params)
    out = {"velocity": []}
    assert vp.compute_on_nodes_(out, cache, []) is out and out["velocity"] == []
    cache.close()
    R = math.pi / 3
    small = vp.filaments.init(vp.define_curve(vp.Ring(), scale=0.5, translate=(1.0, 2.0, 3.0)), 5)
    big = vp.filaments.init(vp.define_curve(vp.Ring(), scale=R, translate=math.pi), 96)
    _check(vp, O, params, [small])
    _check(vp, O, params, [small, big])
