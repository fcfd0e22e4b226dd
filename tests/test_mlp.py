"""The learned correction's network: CPU restatement pinned by finite differences, and the HIP
forward / backward / Adam / BF16-MFMA inference kernels checked against it.

The network has no counterpart in the reference; its arithmetic is defined in
include/vpbs.h and restated in oracle/mlp_oracle.py.
"""
import importlib.util
import pathlib

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def M():
    spec = importlib.util.spec_from_file_location("mlp_oracle", ROOT / "oracle" / "mlp_oracle.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _problem(M, n_in, H, nl, n_out, B, seed=1):
    rng = np.random.default_rng(seed)
    dims = M.layer_dims(n_in, H, nl, n_out)
    p = M.init_params(n_in, H, nl, n_out, seed=seed) + 0.01 * rng.normal(size=M.num_params(n_in, H, nl, n_out))
    X = rng.normal(size=(B, n_in))
    T = rng.normal(size=(B, n_out))
    norm = (rng.normal(size=n_in) * 0.1, 1 + 0.2 * rng.random(n_in), 0.5 + rng.random(n_out))
    return dims, p, X, T, norm


def _fd_grad(M, dims, p, X, T, norm, h=1e-6):
    g = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        lp = M.loss_and_grad(p + e, dims, X, T, *norm)[0]
        lm = M.loss_and_grad(p - e, dims, X, T, *norm)[0]
        g[k] = (lp - lm) / (2 * h)
    return g


def test_backprop_matches_finite_differences(M):
    dims, p, X, T, norm = _problem(M, 8, 6, 2, 3, 20)
    loss, g = M.loss_and_grad(p, dims, X, T, *norm)
    fd = _fd_grad(M, dims, p, X, T, norm)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert err < 1e-7, err
    # the loss is the mean squared normalised error
    y, yn = M.forward(p, dims, X, *norm)
    assert abs(loss - np.mean(np.sum((yn - T / norm[2]) ** 2, axis=1))) < 1e-12 * loss
    assert np.allclose(y, yn * norm[2])


def test_planted_mistake_silu_derivative(M):
    """Backpropagating through σ(z) instead of silu′(z) must fail the finite-difference check."""
    dims, p, X, T, norm = _problem(M, 8, 6, 2, 3, 20)
    _, g = M.loss_and_grad(p, dims, X, T, *norm, dsilu_fn=lambda z: 1 / (1 + np.exp(-z)))
    fd = _fd_grad(M, dims, p, X, T, norm)
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) > 1e-2


def test_adam_first_steps(M):
    """Kingma & Ba: with zero moments the first bias-corrected step is lr·g/(|g| + ε) ≈ lr·sign(g);
    on a quadratic Adam converges."""
    g = np.array([3.0, -0.5, 1e-3])
    p1, m, v = M.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.1)
    assert np.allclose(p1, -0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    p, m, v = np.array([5.0]), np.zeros(1), np.zeros(1)
    for t in range(1, 3001):
        p, m, v = M.adam_step(p, 2 * (p - 1.5), m, v, t, lr=0.05)
    assert abs(p[0] - 1.5) < 1e-3


def test_bf16_rounding(M):
    x = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 3.0e-3, 65504.0], dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1 + 2.0 ** -6  # 1 + 3·2^-8 rounds up to the next bf16 (1 + 4·2^-8)
    assert r[3] == -2.5
    assert abs(r[4] - 3e-3) < 3e-3 * 2.0 ** -8
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


# ----------------------------------------------------------------------------
# GPU: the HIP kernels against the restatement


def _gpu_problem(vp, M, n_in, H, nl, n_out, B, seed=3, n_used=None):
    dims, p, X, T, norm = _problem(M, n_in, H, nl, n_out, B, seed)
    if n_used is not None:
        X[:, n_used:] = 0.0
    net = vp._lib.MLP(n_in, H, nl, n_out)
    net.set_params(p, *norm)
    pf = net.get_params()[0].astype(np.float64)  # FP32 parameters as held by the library
    normf = tuple(np.float32(a).astype(np.float64) for a in norm)
    Xf = X.astype(np.float32)
    Tf = T.astype(np.float32)
    Xd = vp._lib.DeviceArray.from_numpy(Xf)
    Td = vp._lib.DeviceArray.from_numpy(Tf)
    return dims, pf, Xf.astype(np.float64), Tf.astype(np.float64), normf, net, Xd, Td


@pytest.mark.gpu
def test_forward_and_gradient_vs_restatement(vp, M):
    dims, p, X, T, norm, net, Xd, Td = _gpu_problem(vp, M, 128, 128, 3, 3, 1500)
    y = net.forward(Xd).numpy()
    want = M.forward(p, dims, X, *norm)[0]
    ef = np.linalg.norm(y - want) / np.linalg.norm(want)
    loss, g = net.gradient(Xd, Td)
    lw, gw = M.loss_and_grad(p, dims, X, T, *norm)
    eg = np.linalg.norm(g - gw) / np.linalg.norm(gw)
    print(f"  forward {ef:.2e}, loss {abs(loss - lw) / lw:.2e}, gradient {eg:.2e}")
    assert ef < 1e-5
    assert abs(loss - lw) < 1e-5 * lw
    assert eg < 1e-4
    # per layer (the small output layer and the first layer's weights separately)
    off = 0
    for o, i in dims:
        for n in (o * i, o):
            a, b = g[off : off + n], gw[off : off + n]
            assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b) + 1e-9, (o, i, n)
            off += n
    net.close()


@pytest.mark.gpu
def test_adam_step_vs_restatement(vp, M):
    dims, p, X, T, norm, net, Xd, Td = _gpu_problem(vp, M, 32, 32, 2, 3, 700)
    lr = 1e-3
    losses = net.train(Xd, Td, steps=2, batch=10_000, lr=lr, seed=0)  # batch ≥ n: full batch, in order
    pw, m, v = p, np.zeros_like(p), np.zeros_like(p)
    lw = []
    for t in (1, 2):
        l, g = M.loss_and_grad(pw, dims, X, T, *norm)
        lw.append(l)
        pw, m, v = M.adam_step(pw, g, m, v, t, lr)
    got = net.get_params()[0].astype(np.float64)
    step = np.linalg.norm(pw - p)
    print(f"  Adam: losses {losses} vs {lw}; parameter error {np.linalg.norm(got - pw) / step:.2e} of the step")
    assert np.allclose(losses, lw, rtol=1e-5)
    assert np.linalg.norm(got - pw) < 1e-3 * step
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("shape,n", [((128, 128, 3), 2000), ((128, 128, 3), 17), ((128, 128, 2), 1),
                                     ((32, 32, 2), 1000)])
def test_bf16_inference_vs_restatement(vp, M, shape, n):
    n_in, H, nl = shape
    dims, p, X, T, norm, net, Xd, Td = _gpu_problem(vp, M, n_in, H, nl, 3, n, n_used=n_in - 26)
    y = net.infer(Xd).numpy()
    emu = M.infer_bf16(p, dims, X, *norm)
    f64 = M.forward(p, dims, X, *norm)[0]
    e_emu = np.linalg.norm(y - emu) / np.linalg.norm(emu)
    e_f64 = np.linalg.norm(y - f64) / np.linalg.norm(f64)
    print(f"  BF16 MFMA vs emulated BF16 {e_emu:.2e}, vs float64 {e_f64:.2e}")
    assert e_emu < 2e-3
    assert e_f64 < 3e-2
    net.close()


@pytest.mark.gpu
def test_training_reduces_loss(vp, M):
    """Adam on random batches fits a smooth target: the loss drops by more than 20×."""
    rng = np.random.default_rng(5)
    n, n_in = 20000, 32
    X = rng.normal(size=(n, n_in)).astype(np.float32)
    Wt = rng.normal(size=(3, n_in)) / np.sqrt(n_in)
    T = (np.tanh(X @ Wt.T) + 0.3 * np.sin(X[:, :3])).astype(np.float32)
    net = vp._lib.MLP(n_in, 32, 2, 3)
    net.set_params(M.init_params(n_in, 32, 2, 3, seed=2), np.zeros(n_in), np.ones(n_in), np.ones(3))
    Xd, Td = vp._lib.DeviceArray.from_numpy(X), vp._lib.DeviceArray.from_numpy(T)
    l0, _ = net.gradient(Xd, Td)
    net.train(Xd, Td, steps=1500, batch=1024, lr=3e-3, seed=11, record_losses=False)
    l1, _ = net.gradient(Xd, Td)
    print(f"  loss {l0:.3e} -> {l1:.3e}")
    assert l1 < l0 / 20
    net.close()


@pytest.mark.gpu
def test_bf16_inference_output_bound(vp, M):
    """vpbs_mlp_set_output_bound: the BF16 inference clamps each output to ±bound; no bound again
    gives the unclamped outputs back; a non-positive bound is rejected."""
    dims, p, X, T, norm, net, Xd, Td = _gpu_problem(vp, M, 128, 128, 3, 3, 500, n_used=102)
    y0 = net.infer(Xd).numpy()
    b = np.quantile(np.abs(y0), 0.7, axis=0).astype(np.float32)
    net.set_output_bound(b)
    y1 = net.infer(Xd).numpy()
    assert np.array_equal(y1, np.clip(y0, -b, b))
    assert (np.abs(y0) > b).mean() > 0.2
    net.set_output_bound(None)
    assert np.array_equal(net.infer(Xd).numpy(), y0)
    with pytest.raises(vp._lib.VpbsError, match="positive"):
        net.set_output_bound(np.array([1.0, 0.0, 1.0]))
    net.close()
