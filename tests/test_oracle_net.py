"""Pins the FP64 oracle of the learned network (CPU): the convolution against the direct sum
over taps, the pooling/upsampling adjoints by <A x, y> = <x, A^T y>, and the whole backward pass
against central differences of the loss (the pattern of the reference's ml/test_ad.py:7-22:
h = 1e-6 in FP64, at most 40 entries per tensor, relative error |num - ana| / (|num| + |ana| + 1e-8))."""
import numpy as np
import pytest

from oracle import net_oracle as N


def test_conv3_is_the_sum_over_taps():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 3, 5, 8))
    cin, cout = 8, 5
    cp, kp = N.padded(cin, cout)
    W = np.zeros((cp, kp)); W[:cout, :27 * cin] = rng.standard_normal((cout, 27 * cin))
    b = np.zeros(cp); b[:cout] = rng.standard_normal(cout)
    y, _ = N.conv3(x, W, b, cout, relu=False)
    ref = np.zeros((2, 4, 3, 5, cout)) + b[:cout]
    for t in range(27):
        di, dj, dk = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        Wt = W[:cout, t * cin:(t + 1) * cin]
        for i in range(4):
            for j in range(3):
                for k in range(5):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < 4 and 0 <= jj < 3 and 0 <= kk < 5:
                        ref[:, i, j, k] += x[:, ii, jj, kk] @ Wt.T
    assert np.abs(y - ref).max() < 1e-12


def test_pool_and_up_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 6, 2, 3))
    yc = rng.standard_normal((1, 2, 3, 1, 3))
    assert abs((N.pool(x) * yc).sum() - (x * N.pool_bwd(yc)).sum()) < 1e-12
    assert abs((N.up(yc) * x).sum() - (yc * N.up_bwd(x)).sum()) < 1e-12
    assert np.allclose(N.pool(N.up(yc)), yc)


def test_im2col_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 3, 4, 2, 8))
    c = rng.standard_normal((1, 3, 4, 2, 27 * 8))
    assert abs((N.im2col(x) * c).sum() - (x * N.col2im(c, 8)).sum()) < 1e-11


def test_backward_matches_central_differences():
    """Every parameter tensor -- 10 weights at seeded random positions of the layer's live block,
    its 6 largest analytic weight gradients (the coarsest level's taps mostly fall in the zero
    padding, so random entries there are exactly zero) and 6 biases -- and 40 input features, on a
    16x8x8 raster with positive biases (no dead layers)."""
    rng = np.random.default_rng(0)
    P = N.init_params(seed=3)
    for name in P:
        W, b = P[name]
        cout = next(l for l in N.LAYERS if l[0] == name)[2]
        b[:cout] = 0.1 + 0.05 * rng.standard_normal(cout)
    shape = (1, 16, 8, 8)
    f = rng.standard_normal(shape + (N.IN_CH,))
    t = rng.standard_normal(shape + (3,))
    mask = rng.uniform(size=shape) < 0.8

    def L():
        out, _ = N.forward(P, f)
        return N.loss(out, t, mask)[0]
    out, C = N.forward(P, f)
    _, dout = N.loss(out, t, mask)
    G, df = N.backward(P, C, dout)
    h = 1e-6
    worst, checked = 0.0, 0
    for name, cin, cout, _ in N.LAYERS:
        W, b = P[name]
        dW, db = G[name]
        live = np.abs(dW[:cout, :27 * cin])
        top = np.argsort(live.ravel())[-6:]
        picks = [("W", (int(rng.integers(cout)), int(rng.integers(27 * cin)))) for _ in range(10)]
        picks += [("W", tuple(int(q) for q in np.unravel_index(i, live.shape))) for i in top]
        picks += [("b", (int(q),)) for q in rng.choice(cout, min(cout, 6), replace=False)]
        assert live.max() > 0, name
        for kind, idx in picks:
            arr, ana = (W, dW[idx]) if kind == "W" else (b, db[idx])
            old = arr[idx]
            arr[idx] = old + h; lp = L()
            arr[idx] = old - h; lm = L()
            arr[idx] = old
            num = (lp - lm) / (2 * h)
            worst = max(worst, abs(num - ana) / (abs(num) + abs(ana) + 1e-8))
            checked += 1
    for _ in range(40):
        idx = tuple(int(rng.integers(s)) for s in f.shape)
        old = f[idx]
        f[idx] = old + h; lp = L()
        f[idx] = old - h; lm = L()
        f[idx] = old
        num = (lp - lm) / (2 * h)
        worst = max(worst, abs(num - df[idx]) / (abs(num) + abs(df[idx]) + 1e-8))
    assert checked == sum(16 + min(cout, 6) for _, _, cout, _ in N.LAYERS) and worst < 1e-3, worst
