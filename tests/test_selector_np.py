"""The numpy restatement of the selector's MLP that the GPU tests check the
HIP kernels against, pinned by finite differences of its loss, by a closed
form (a one-hidden-unit net reduces to a known function), and by Adam's
first step (a step of exactly lr against the gradient's sign)."""
import numpy as np

from tests import selector_np as S


def _weights(seed):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.3, S.NP + 2 * S.NF)
    w[S.NP:S.NP + S.NF] = rng.normal(0, 1, S.NF)        # mean
    w[S.NP + S.NF:] = rng.uniform(0.5, 2.0, S.NF)       # 1/std
    return w


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = _weights(2)
    X = rng.normal(0, 1, (7, S.NF))
    Y = rng.normal(0, 1, (7, S.NO))
    M = (rng.uniform(0, 1, (7, S.NO)) > 0.3).astype(float)
    _, g = S.loss_grad(w, X, Y, M)
    h = 1e-6
    for k in rng.choice(S.NP, 60, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (S.loss_grad(wp, X, Y, M)[0] - S.loss_grad(wm, X, Y, M)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_forward_closed_form():
    """Only hidden unit 0 of each layer and output 0 non-zero: y0 =
    b3 + c tanh(b + a tanh(u x_k + d)), the other outputs their biases."""
    w = np.zeros(S.NP + 2 * S.NF)
    W1, b1, W2, b2, W3, b3, mu, isd = S.unpack(np.arange(w.size))  # index views
    idx = np.arange(w.size)
    o = 0
    W1i = idx[o:o + S.NH * S.NF].reshape(S.NH, S.NF); o += S.NH * S.NF
    b1i = idx[o:o + S.NH]; o += S.NH
    W2i = idx[o:o + S.NH * S.NH].reshape(S.NH, S.NH); o += S.NH * S.NH
    b2i = idx[o:o + S.NH]; o += S.NH
    W3i = idx[o:o + S.NO * S.NH].reshape(S.NO, S.NH); o += S.NO * S.NH
    b3i = idx[o:o + S.NO]; o += S.NO
    mui = idx[o:o + S.NF]; o += S.NF
    isdi = idx[o:o + S.NF]
    u, d, a, b, c = 0.7, -0.2, 1.3, 0.1, 2.0
    w[isdi] = 1.0
    w[mui[3]] = 0.5
    w[W1i[0, 3]] = u
    w[b1i[0]] = d
    w[W2i[0, 0]] = a
    w[b2i[0]] = b
    w[W3i[0, 0]] = c
    w[b3i] = [0.3, -1.0, 4.0]
    X = np.linspace(-2, 2, 9)[:, None] * np.ones((1, S.NF))
    y, _ = S.forward(w, X)
    h1_0 = np.tanh(u * (X[:, 3] - 0.5) + d)
    # the other first-layer units are tanh(0) = 0, so layer 2 sees only unit 0
    want = 0.3 + c * np.tanh(b + a * h1_0)
    assert np.allclose(y[:, 0], want, rtol=0, atol=1e-14)
    assert np.allclose(y[:, 1], -1.0) and np.allclose(y[:, 2], 4.0)


def test_adam_first_step_is_lr_times_sign():
    rng = np.random.default_rng(3)
    w = _weights(4)
    X = rng.normal(0, 1, (5, S.NF))
    Y = rng.normal(0, 1, (5, S.NO))
    _, g = S.loss_grad(w, X, Y)
    w1, _ = S.adam(w, X, Y, steps=1, lr=1e-3, eps=0.0)
    step = w1[:S.NP] - w[:S.NP]
    nz = np.abs(g) > 1e-12
    assert np.allclose(step[nz], -1e-3 * np.sign(g[nz]), rtol=1e-9, atol=0)
    assert np.array_equal(w1[S.NP:], w[S.NP:])


def test_choose_cheapest_passing():
    y = np.array([[-7, -5, -2], [-7, -5, -4.5], [-7, -3, -2], [-3, -2, -1]])
    assert list(S.choose(y, 1e-4)) == [4, 2, 1, 0]
    assert list(S.choose(y, 1e-4, mask=(1,))) == [1, 1, 1, 0]
