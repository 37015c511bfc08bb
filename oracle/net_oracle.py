"""FP64 CPU oracle of the learned warm-start network (numpy).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  The product never imports this module.

The network (comech_amd.learned.net has the same layers): a 4-level 3-D U-Net of 3x3x3 convolutions
with zero padding, 16/32/64/128 channels, two convolutions per level, 2x2x2 average pooling down,
nearest upsampling up, pre-activation additive skips (y = relu(conv(up) + skip)), and a linear
3-channel head.  No counterpart in the reference (its solve path has no learned part); the
gradient check follows the reference's own autodiff test pattern (ml/test_ad.py:7-22: central
differences with h = 1e-6 in FP64 on at most 40 entries per input, loss = sum(y * cos(index))).

Weights use the product's layout: W[co][kp] (Coutp rows, Kp columns, zero padding), kp = tap Cin +
ci, tap = 9 (di + 1) + 3 (dj + 1) + (dk + 1).

How each part is pinned (tests/test_oracle_net.py): ``conv3`` against the direct sum over taps
at every voxel of a small raster; ``pool``/``up`` and their adjoints by <A x, y> = <x, A^T y>;
``backward`` against central differences of ``loss``.
"""
import numpy as np

IN_CH = 16
# name, Cin, Cout, ReLU; the order of the forward pass
LAYERS = [("e0a", 16, 16, True), ("e0b", 16, 16, True),
          ("e1a", 16, 32, True), ("e1b", 32, 32, True),
          ("e2a", 32, 64, True), ("e2b", 64, 64, True),
          ("e3a", 64, 128, True), ("e3b", 128, 128, True),
          ("d2a", 128, 64, True), ("d2b", 64, 64, True),
          ("d1a", 64, 32, True), ("d1b", 32, 32, True),
          ("d0a", 32, 16, True), ("d0b", 16, 16, True),
          ("head", 16, 3, False)]


def padded(cin, cout):
    """(Coutp, Kp): outputs rounded up to 16, 27 Cin rounded up to 32."""
    return ((cout + 15) // 16) * 16, ((27 * cin + 31) // 32) * 32


def init_params(seed=0):
    """He-normal weights (std sqrt(2 / (27 Cin))), the head scaled by 0.1; zero biases."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, cin, cout, _ in LAYERS:
        cp, kp = padded(cin, cout)
        W = np.zeros((cp, kp))
        std = np.sqrt(2.0 / (27 * cin)) * (0.1 if name == "head" else 1.0)
        W[:cout, :27 * cin] = rng.standard_normal((cout, 27 * cin)) * std
        P[name] = (W, np.zeros(cp))
    return P


def im2col(x):
    """x [B, nx, ny, nz, C] -> [B, nx, ny, nz, 27 C], column tap * C + c, zero padding."""
    B, nx, ny, nz, C = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    return np.concatenate([xp[:, t // 9:t // 9 + nx, (t // 3) % 3:(t // 3) % 3 + ny, t % 3:t % 3 + nz, :]
                           for t in range(27)], -1)


def col2im(col, C):
    """Adjoint of im2col."""
    B, nx, ny, nz, _ = col.shape
    xp = np.zeros((B, nx + 2, ny + 2, nz + 2, C))
    for t in range(27):
        xp[:, t // 9:t // 9 + nx, (t // 3) % 3:(t // 3) % 3 + ny, t % 3:t % 3 + nz, :] += col[..., t * C:(t + 1) * C]
    return xp[:, 1:-1, 1:-1, 1:-1, :]


def conv3(x, W, b, cout, add=None, relu=True):
    """y = act(sum_taps W x + b (+ add)); returns (y, cache)."""
    cin = x.shape[-1]
    col = im2col(x)
    pre = col @ W[:cout, :27 * cin].T + b[:cout]
    if add is not None:
        pre = pre + add
    y = np.maximum(pre, 0.0) if relu else pre
    return y, (col, pre, relu, cin, cout)


def conv3_bwd(dy, W, cache):
    """(dx, dW (padded layout), db (padded), dadd) of conv3."""
    col, pre, relu, cin, cout = cache
    g = dy * (pre > 0) if relu else dy
    g2 = g.reshape(-1, cout)
    dW = np.zeros_like(W)
    dW[:cout, :27 * cin] = g2.T @ col.reshape(-1, 27 * cin)
    db = np.zeros(W.shape[0])
    db[:cout] = g2.sum(0)
    dx = col2im(g @ W[:cout, :27 * cin], cin)
    return dx, dW, db, g


def pool(x):
    B, nx, ny, nz, C = x.shape
    return x.reshape(B, nx // 2, 2, ny // 2, 2, nz // 2, 2, C).mean((2, 4, 6))


def pool_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(dy):
    B, nx, ny, nz, C = dy.shape
    return dy.reshape(B, nx // 2, 2, ny // 2, 2, nz // 2, 2, C).sum((2, 4, 6))


def forward(P, f):
    """f [B, nx, ny, nz, 16] (sizes divisible by 8) -> (out [B, nx, ny, nz, 3], caches)."""
    C = {}
    h = {}

    def cv(name, x, add=None):
        W, b = P[name]
        spec = next(l for l in LAYERS if l[0] == name)
        y, C[name] = conv3(x, W, b, spec[2], add=add, relu=spec[3])
        return y
    h["e0a"] = cv("e0a", f); h["e0"] = cv("e0b", h["e0a"])
    h["e1a"] = cv("e1a", pool(h["e0"])); h["e1"] = cv("e1b", h["e1a"])
    h["e2a"] = cv("e2a", pool(h["e1"])); h["e2"] = cv("e2b", h["e2a"])
    h["e3a"] = cv("e3a", pool(h["e2"])); h["e3"] = cv("e3b", h["e3a"])
    h["d2a"] = cv("d2a", up(h["e3"]), add=h["e2"]); h["d2"] = cv("d2b", h["d2a"])
    h["d1a"] = cv("d1a", up(h["d2"]), add=h["e1"]); h["d1"] = cv("d1b", h["d1a"])
    h["d0a"] = cv("d0a", up(h["d1"]), add=h["e0"]); h["d0"] = cv("d0b", h["d0a"])
    out = cv("head", h["d0"])
    return out, C


def loss(out, t, mask):
    """Mean over the masked cells and the 3 channels of (out - t)^2, and d loss / d out."""
    w = mask[..., None].astype(float)
    n = max(1.0, 3.0 * mask.sum())
    e = out - t
    return float((w * e * e).sum() / n), 2.0 * w * e / n


def backward(P, C, dout):
    """Gradients {name: (dW, db)} of the network for d loss / d out."""
    G = {}

    def bw(name, dy):
        dx, dW, db, g = conv3_bwd(dy, P[name][0], C[name])
        G[name] = (dW, db)
        return dx, g
    d_d0, _ = bw("head", dout)
    d_d0a, _ = bw("d0b", d_d0)
    d_u0, g = bw("d0a", d_d0a)
    d_e0 = g.copy()                         # skip: d pre / d e0 = 1
    d_d1 = up_bwd(d_u0)
    d_d1a, _ = bw("d1b", d_d1)
    d_u1, g = bw("d1a", d_d1a)
    d_e1 = g.copy()
    d_d2 = up_bwd(d_u1)
    d_d2a, _ = bw("d2b", d_d2)
    d_u2, g = bw("d2a", d_d2a)
    d_e2 = g.copy()
    d_e3 = up_bwd(d_u2)
    d_e3a, _ = bw("e3b", d_e3)
    d_p3, _ = bw("e3a", d_e3a)
    d_e2 = d_e2 + pool_bwd(d_p3)
    d_e2a, _ = bw("e2b", d_e2)
    d_p2, _ = bw("e2a", d_e2a)
    d_e1 = d_e1 + pool_bwd(d_p2)
    d_e1a, _ = bw("e1b", d_e1)
    d_p1, _ = bw("e1a", d_e1a)
    d_e0 = d_e0 + pool_bwd(d_p1)
    d_e0a, _ = bw("e0b", d_e0)
    d_f, _ = bw("e0a", d_e0a)
    return G, d_f
