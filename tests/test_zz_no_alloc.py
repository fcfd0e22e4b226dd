"""vpbs_energy_spectrum_noalloc (partial spectra in the library's fixed device array, no
allocation) against vpbs_energy_spectrum (per-call hipMallocAsync buffer) on the same kept
field: FP64 spectra agree to 1e-12 relative (the two use different block counts, so the
summation order differs); two calls in a row, the second on a smaller field with fewer bins,
then a third on the first field again, into device memory.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TWO_PI = 2 * math.pi


def _kept(vp, Ns, nrings, seed):
    fs = vp.tangles.random_rings(nrings, 0.5, 1.5, TWO_PI / Ns, seed=seed)
    spl = vp.GaussianSplitting.create(Ls=TWO_PI, β=3.5, Ns=Ns)
    p = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, Δ=0.25, splitting=spl)
    ctx = vp._lib.Context(p)
    ctx.set_filaments(vp.FilamentBatch(fs), p)
    ctx.compute(vp._lib.VELOCITY | vp._lib.KEEP_UHAT, fetch=False)
    return ctx


def test_noalloc_spectrum_matches_allocating_one(vp):
    big = _kept(vp, 128, 64, seed=7)
    small = _kept(vp, 48, 6, seed=8)
    for ctx, nk, dk in ((big, 700, 0.2), (small, 30, 1.0)):
        want = ctx.energy_spectrum(nk, dk)
        got = ctx.energy_spectrum_noalloc(nk, dk)
        err = rel_l2(got, want)
        print(f"  nk = {nk}: no-alloc vs allocating {err:.2e}; E sum {want.sum():.6e}", flush=True)
        assert np.count_nonzero(want) > nk // 3
        assert err <= 1e-12
    # first field again, output in device memory
    out = vp._lib.DeviceArray.from_numpy(np.full(700, np.nan))
    big.energy_spectrum_noalloc(700, 0.2, out=out)
    assert rel_l2(out.numpy(), big.energy_spectrum(700, 0.2)) <= 1e-12
    with pytest.raises(vp._lib.VpbsError, match="without allocation"):
        big.energy_spectrum_noalloc(1025, 0.2)
    big.close()
    small.close()
