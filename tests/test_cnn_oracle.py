"""CPU checks of the FP64 oracle of the warm-start network layers
(oracle/cnn.py): brute-force loops of the definition, the adjoint identity
<conv(x), dy> = <x, bwd_data(dy)> and finite differences of the weights."""
import numpy as np

from oracle import cnn as R


def test_conv_matches_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 5, 7))
    w = rng.standard_normal((4, 3, 3, 3))
    bias = rng.standard_normal(4)
    assert np.allclose(R.conv3x3(x, w, bias), R.conv3x3_loops(x, w, bias), rtol=1e-12, atol=1e-12)


def test_bwd_data_is_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 6, 5))
    w = rng.standard_normal((4, 3, 3, 3))
    dy = rng.standard_normal((2, 4, 6, 5))
    lhs = np.sum(R.conv3x3(x, w) * dy)
    rhs = np.sum(x * R.conv3x3_bwd_data(dy, w))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_bwd_weights_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 2, 5, 6))
    w = rng.standard_normal((3, 2, 3, 3))
    bias = rng.standard_normal(3)
    dy = rng.standard_normal((2, 3, 5, 6))
    dw, db = R.conv3x3_bwd_weights(x, dy)
    loss = lambda w_, b_: np.sum(R.conv3x3(x, w_, b_) * dy)
    h = 1e-6
    for idx in [(0, 0, 0, 0), (2, 1, 1, 2), (1, 0, 2, 1)]:
        wp = w.copy(); wp[idx] += h
        wm = w.copy(); wm[idx] -= h
        assert abs((loss(wp, bias) - loss(wm, bias)) / (2 * h) - dw[idx]) < 1e-6 * max(1, abs(dw[idx]))
    bp = bias.copy(); bp[1] += h
    bm = bias.copy(); bm[1] -= h
    assert abs((loss(w, bp) - loss(w, bm)) / (2 * h) - db[1]) < 1e-6 * max(1, abs(db[1]))


def test_bf16_rounding():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even mantissa
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e38], np.float32)
    r = R.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6 and r[3] == -2.5


def test_net_grads_finite_differences():
    # backpropagation of the oracle network against central differences
    rng = np.random.default_rng(5)
    ch = 3
    sizes = R.net_channels(ch)
    n = sum(sizes[l + 1] * sizes[l] * 9 + sizes[l + 1] for l in range(6))
    p = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((2, 2, 5, 4))
    t = rng.standard_normal((2, 2, 5, 4))
    g, loss = R.net_grads(p, x, t, ch, rounding=False)
    h = 1e-6
    for i in rng.choice(n, 12, replace=False):
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (R.net_grads(pp, x, t, ch, False)[1] - R.net_grads(pm, x, t, ch, False)[1]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
