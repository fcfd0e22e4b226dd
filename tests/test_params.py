"""Parameter derivation against the reference's documented known answers."""
import json
import math
import pathlib

import numpy as np
import pytest

KA = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_known_answers.json").read_text())


def test_gaussian_splitting_docstring(vp):
    ka = KA["gaussian_splitting_docstring"]
    s = vp.GaussianSplitting.create(Ls=ka["Ls"], β=ka["beta"], Ns=ka["Ns"])
    assert s.α == ka["alpha"]
    assert s.rcut == ka["rcut"]
    assert s.rcut / min(s.Ls) == ka["rcut_over_Lmin"]
    assert s.kmax == ka["kmax"]
    assert s.rtol == ka["rtol"]
    assert s.β_shortrange == pytest.approx(3.5, rel=1e-15)
    assert s.β_longrange == pytest.approx(3.5, rel=1e-15)


@pytest.mark.parametrize("beta,rtol", KA["accuracy_table"]["beta_to_rtol"])
def test_accuracy_model(vp, beta, rtol):
    s = vp.GaussianSplitting.create(Ls=2 * math.pi, β=beta, α=5.0)
    assert s.rtol == pytest.approx(rtol, rel=2e-3)


def test_splitting_constructions_consistent(vp):
    L = 2 * math.pi
    a = vp.GaussianSplitting.create(Ls=L, β=3.5, α=9.0)
    b = vp.GaussianSplitting.create(Ls=L, β=3.5, rcut=a.rcut)
    c = vp.GaussianSplitting.create(Ls=L, rtol=a.rtol, α=9.0)
    assert a.Ns == b.Ns == c.Ns
    assert a.α == pytest.approx(b.α) and a.rcut == pytest.approx(c.rcut)
    # kmax_to_gridsize: kmax = 63 with L = 2π needs m = 63 -> N = 128
    assert vp.params.kmax_to_gridsize(63.0, L) == 128
    # non-uniform periods (test/trefoil.jl:395-396)
    d = vp.GaussianSplitting.create(Ls=(1.5 * math.pi, 1.5 * math.pi, 2 * math.pi), rtol=1e-6, α=3.0)
    assert d.Ns[0] == d.Ns[1] and d.Ns[2] > d.Ns[0]


def test_cell_list_docstring(O):
    ka = KA["cell_list_docstring"]
    Nc = np.zeros(3, dtype=np.int32)
    Ls = np.full(3, ka["Ls"])
    O.lib().or_cell_dims(ka["rcut"], O._p(Ls), ka["nsubdiv"], O._p(Nc, O._ip))
    assert list(Nc) == [ka["ncells"]] * 3
    assert ka["Ls"] / Nc[0] == ka["cell_size"]
    assert ka["Ls"] / Nc[0] * ka["nsubdiv"] == ka["effective_rcut"]


def test_parameter_validation(vp):
    L = 2 * math.pi
    # rcut above the cell-list limit M/(2M+1)·L (params.jl:99-104)
    s = vp.GaussianSplitting.create(Ls=L, α=1.0, rcut=0.45 * L, Ns=16)
    with pytest.raises(ValueError, match="larger than that allowed"):
        vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=s)
    # with nsubdiv = 4 the limit is 4/9 L < 0.45 L; with a naive-like bound the check is 2rcut ≤ L
    with pytest.raises(ValueError):
        vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, splitting=s,
                                  backend_short=vp.HIPShortRangeBackend(4))
    with pytest.raises(ValueError):
        vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6)  # no splitting
    with pytest.raises(TypeError):
        vp.ParamsBiotSavart.build(np.int32, Γ=1.0, a=1e-6, splitting=vp.GaussianSplitting.create(Ls=L, β=3.5, Ns=64))
    p = vp.ParamsBiotSavart.build(np.float32, Γ=2.0, a=1e-6, splitting=vp.GaussianSplitting.create(Ls=L, β=3.5, Ns=64))
    assert p.precision == "f32" and p.rcut_sq == p.rcut * p.rcut
    assert len(p.quadrature) == 3 and p.backend_short.nsubdiv == 2 and p.backend_long.m == 4


def test_deprecated_keywords(vp):
    p = vp.ParamsBiotSavart.build(np.float64, Γ=1.0, a=1e-6, α=9.0, Ls=2 * math.pi, Ns=128, rcut=0.38)
    assert p.α == 9.0 and p.rcut == 0.38 and p.Ns == (128, 128, 128)
    assert p.splitting.rtol == pytest.approx(10 ** -1.3 * math.exp(-(9 * 0.38) ** 2))
