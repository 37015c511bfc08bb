"""The learned warm start's network on the GPU: a 4-level 3-D U-Net (16/32/64/128 channels,
3x3x3 convolutions, 2x2x2 average pooling, nearest upsampling, pre-activation additive skips,
linear 3-channel head) whose every product runs on the BF16 matrix cores with FP32 accumulation
(``cm_nn_gemm_tn``, include/comech_learned.h) and whose training (backward pass, Adam on FP32
master weights) runs in the project's own kernels.  torch provides device buffers only.

No counterpart in the reference; the CPU oracle of the same network is oracle/net_oracle.py.

Layout: feature maps channels-last [B * V][C], V the raster's cells in the reference's order
(i * ny + j) * nz + k; sizes divisible by 8 (three poolings).  A layer keeps FP32 master weights
W [Coutp][Kp] (kp = tap * Cin + ci), biases b [Coutp], their Adam moments and gradients, and BF16
copies Wt (forward) and Wf (flipped transpose, input gradient).
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from .._lib import check_nn

IN_CH = 16
OUT_CH = 3
LAYERS = [("e0a", 16, 16, True), ("e0b", 16, 16, True),
          ("e1a", 16, 32, True), ("e1b", 32, 32, True),
          ("e2a", 32, 64, True), ("e2b", 64, 64, True),
          ("e3a", 64, 128, True), ("e3b", 128, 128, True),
          ("d2a", 128, 64, True), ("d2b", 64, 64, True),
          ("d1a", 64, 32, True), ("d1b", 32, 32, True),
          ("d0a", 32, 16, True), ("d0b", 16, 16, True),
          ("head", 16, OUT_CH, False)]
SPEC = {name: (cin, cout, relu) for name, cin, cout, relu in LAYERS}
EPI_F32, EPI_BF16, EPI_ATOMIC = 0, 1, 2


def padded(cin, cout):
    """(Coutp, Kp): outputs rounded up to 16, 27 Cin rounded up to 32."""
    return ((cout + 15) // 16) * 16, ((27 * cin + 31) // 32) * 32


def _vp(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def to_bf16(a):
    """numpy float array -> BF16 bit patterns (int16 view), round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16).view(np.int16)


def from_bf16(b):
    """BF16 bit patterns (int16/uint16 numpy) -> float32."""
    return (np.asarray(b).view(np.uint16).astype(np.uint32) << 16).view(np.float32)


class Net:
    """Parameters, optimizer state and work buffers of the network on one device."""

    def __init__(self, device, seed=0, params=None):
        self.dev = torch.device(device)
        self.lib = _lib.load()
        self.layers = {}
        rng = np.random.default_rng(seed)
        for name, cin, cout, relu in LAYERS:
            cp, kp = padded(cin, cout)
            kpf = padded(cp, cin)[1]
            if params is not None:
                W, b = params[name]
            else:                          # He normal (std sqrt(2 / fan_in)); the head scaled by 0.1
                W = np.zeros((cp, kp))
                std = np.sqrt(2.0 / (27 * cin)) * (0.1 if name == "head" else 1.0)
                W[:cout, :27 * cin] = rng.standard_normal((cout, 27 * cin)) * std
                b = np.zeros(cp)
            f32 = lambda a: torch.tensor(np.asarray(a, np.float32), device=self.dev)
            L = dict(cin=cin, cout=cout, cp=cp, kp=kp, kpf=kpf, relu=relu, W=f32(W), b=f32(b))
            for k in ("mW", "vW", "gW"):
                L[k] = torch.zeros_like(L["W"])
            for k in ("mb", "vb", "gb"):
                L[k] = torch.zeros_like(L["b"])
            L["Wt"] = torch.empty((cp, kp), dtype=torch.int16, device=self.dev)
            L["Wf"] = torch.empty((cin, kpf), dtype=torch.int16, device=self.dev)
            self.layers[name] = L
        self.step_count = 0
        self._ws = {}
        self.refresh()

    # ------------------------------------------------------------------ parameters
    @property
    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)

    def refresh(self):
        """BF16 compute copies of every layer's master weights."""
        for L in self.layers.values():
            check_nn(self.lib.cm_nn_weights(self.stream, L["cin"], L["cp"], L["kp"], L["kpf"], _vp(L["W"]),
                                            _vp(L["Wt"]), _vp(L["Wf"])))

    def params(self):
        """{name: (W [Coutp][Kp], b [Coutp])} as FP64 numpy arrays (the oracle's layout)."""
        return {n: (L["W"].double().cpu().numpy(), L["b"].double().cpu().numpy()) for n, L in self.layers.items()}

    def grads(self):
        return {n: (L["gW"].double().cpu().numpy(), L["gb"].double().cpu().numpy()) for n, L in self.layers.items()}

    def save(self, path, **meta):
        arrs = {}
        for n, L in self.layers.items():
            arrs[n + ".W"] = L["W"].cpu().numpy()
            arrs[n + ".b"] = L["b"].cpu().numpy()
        for k, v in meta.items():
            arrs["meta." + k] = np.asarray(v)
        np.savez_compressed(path, **arrs)

    @classmethod
    def load(cls, path, device):
        z = np.load(path)
        params = {name: (z[name + ".W"], z[name + ".b"]) for name, *_ in LAYERS}
        net = cls(device, params=params)
        net.meta = {k[5:]: z[k] for k in z.files if k.startswith("meta.")}
        return net

    # ------------------------------------------------------------------ buffers
    def _buf(self, key, n, dtype):
        t = self._ws.get(key)
        if t is None or t.numel() < n or t.dtype != dtype:
            t = torch.empty(max(n, 1), dtype=dtype, device=self.dev)
            self._ws[key] = t
        return t[:n]

    # ------------------------------------------------------------------ layers
    def _conv(self, name, x, dims, B, out, add=None, f32_out=False):
        """out = act(conv(x) + b (+ add)); x BF16 [M][cin], out BF16 (or FP32) [M][Coutp]."""
        L = self.layers[name]
        M = B * dims[0] * dims[1] * dims[2]
        col = self._buf("col", M * L["kp"], torch.int16)
        check_nn(self.lib.cm_nn_im2col(self.stream, B, *dims, L["cin"], L["kp"], _vp(x), _vp(col), 0, 0))
        check_nn(self.lib.cm_nn_gemm_tn(self.stream, M, L["cp"], L["kp"], _vp(col), L["kp"], _vp(L["Wt"]), L["kp"],
                                        _vp(out), L["cp"], _vp(L["b"]), _vp(add), L["cp"] if add is not None else 0,
                                        int(L["relu"]), EPI_F32 if f32_out else EPI_BF16, 1, 1.0))
        return out

    def _conv_bwd(self, name, x, y, dy, dims, B, dx, dskip=None):
        """Gradients of out = act(conv(x) + b (+ skip)) for dy (FP32 [M][Coutp]): accumulates the
        layer's gW, gb, dx (FP32 [M][cin], if not None) and dskip (FP32 [M][Coutp], if not None)."""
        L = self.layers[name]
        M = B * dims[0] * dims[1] * dims[2]
        Mp = (M + 31) // 32 * 32
        cp, kp, cin, kpf = L["cp"], L["kp"], L["cin"], L["kpf"]
        g = self._buf("g", M * cp, torch.int16)
        g32 = self._buf("g32", M * cp, torch.float32)
        check_nn(self.lib.cm_nn_relu_bwd(self.stream, M * cp, _vp(dy), _vp(y) if L["relu"] else None, _vp(g),
                                         _vp(g32)))
        check_nn(self.lib.cm_nn_colsum(self.stream, M, cp, _vp(g32), _vp(L["gb"])))
        if dskip is not None:
            check_nn(self.lib.cm_nn_axpy(self.stream, M * cp, 1.0, _vp(g32), _vp(dskip)))
        # weights: gW[co][kp] += sum_m g[m][co] col[m][kp]  (K = voxels, split over blocks)
        gT = self._buf("gT", cp * Mp, torch.int16)
        check_nn(self.lib.cm_nn_transpose(self.stream, M, Mp, cp, _vp(g), None, _vp(gT)))
        colT = self._buf("colT", kp * Mp, torch.int16)
        check_nn(self.lib.cm_nn_im2col(self.stream, B, *dims, cin, kp, _vp(x), _vp(colT), 1, Mp))
        tiles = ((cp + 63) // 64) * ((kp + 127) // 128)
        ksplit = int(max(1, min(Mp // 32 // 8, -(-1024 // tiles))))
        check_nn(self.lib.cm_nn_gemm_tn(self.stream, cp, kp, Mp, _vp(gT), Mp, _vp(colT), Mp, _vp(L["gW"]), kp, None,
                                        None, 0, 0, EPI_ATOMIC, ksplit, 1.0))
        if dx is not None:
            # input: dx += conv(g, Wf), the flipped transposed weights
            colg = self._buf("col", M * kpf, torch.int16)
            check_nn(self.lib.cm_nn_im2col(self.stream, B, *dims, cp, kpf, _vp(g), _vp(colg), 0, 0))
            check_nn(self.lib.cm_nn_gemm_tn(self.stream, M, cin, kpf, _vp(colg), kpf, _vp(L["Wf"]), kpf, _vp(dx), cin,
                                            None, None, 0, 0, EPI_ATOMIC, 1, 1.0))

    # ------------------------------------------------------------------ network
    def forward(self, F, dims, B=1, keep=False):
        """F: BF16 (int16 storage) features [B V][16], dims (nx, ny, nz) divisible by 8.  Returns
        the FP32 output [B V][16] (first 3 channels used).  keep: hold the activations for
        backward."""
        nx, ny, nz = dims
        if nx % 8 or ny % 8 or nz % 8:
            raise ValueError("raster sizes must be divisible by 8 (three 2x2x2 poolings)")
        d = [(nx >> l, ny >> l, nz >> l) for l in range(4)]
        Ms = [B * a * b * c for a, b, c in d]
        h = {"f": F}
        act = lambda key, l, c: self._buf("a." + key, Ms[l] * c, torch.int16)
        st = self.stream
        h["e0a"] = self._conv("e0a", F, d[0], B, act("e0a", 0, 16))
        h["e0"] = self._conv("e0b", h["e0a"], d[0], B, act("e0", 0, 16))
        for l, src, na, nb_, c_in, c in ((1, "e0", "e1a", "e1b", 16, 32), (2, "e1", "e2a", "e2b", 32, 64),
                                          (3, "e2", "e3a", "e3b", 64, 128)):
            p = act("p%d" % l, l, c_in)
            check_nn(self.lib.cm_nn_pool(st, B, *d[l - 1], c_in, _vp(h[src]), _vp(p)))
            h["p%d" % l] = p
            h[na] = self._conv(na, p, d[l], B, act(na, l, c))
            h["e%d" % l] = self._conv(nb_, h[na], d[l], B, act("e%d" % l, l, c))
        for l, src, na, nb_, c_in, c in ((2, "e3", "d2a", "d2b", 128, 64), (1, "d2", "d1a", "d1b", 64, 32),
                                          (0, "d1", "d0a", "d0b", 32, 16)):
            u = act("u%d" % l, l, c_in)
            check_nn(self.lib.cm_nn_up(st, B, *d[l + 1], c_in, _vp(h[src]), _vp(u)))
            h["u%d" % l] = u
            h[na] = self._conv(na, u, d[l], B, act(na, l, c), add=h["e%d" % l])
            h["d%d" % l] = self._conv(nb_, h[na], d[l], B, act("d%d" % l, l, c))
        out = self._buf("out", Ms[0] * 16, torch.float32)
        self._conv("head", h["d0"], d[0], B, out, f32_out=True)
        if keep:
            self._h, self._dims, self._B = h, d, B
        return out

    def backward(self, dout):
        """Gradients of the last forward(keep=True) for d loss / d out (FP32 [B V][16]) into
        every layer's gW, gb (accumulated: zero_grad first)."""
        h, d, B = self._h, self._dims, self._B
        Ms = [B * a * b * c for a, b, c in d]
        st = self.stream
        gz = lambda key, l, c: self._buf("d." + key, Ms[l] * c, torch.float32).zero_()
        dd0 = gz("d0", 0, 16)
        self._conv_bwd("head", h["d0"], None, dout, d[0], B, dd0)
        de = {0: gz("e0", 0, 16), 1: gz("e1", 1, 32), 2: gz("e2", 2, 64), 3: gz("e3", 3, 128)}
        dsrc = dd0
        for l, na, nb_, c_in, c, up_src in ((0, "d0a", "d0b", 32, 16, "d1"), (1, "d1a", "d1b", 64, 32, "d2"),
                                             (2, "d2a", "d2b", 128, 64, "e3")):
            dna = gz(na, l, c)
            self._conv_bwd(nb_, h[na], h["d%d" % l], dsrc, d[l], B, dna)
            du = gz("u%d" % l, l, c_in)
            self._conv_bwd(na, h["u%d" % l], h[na], dna, d[l], B, du, dskip=de[l])
            if up_src == "e3":
                check_nn(self.lib.cm_nn_up_bwd(st, B, *d[l + 1], c_in, _vp(du), _vp(de[3])))
            else:
                dnext = gz(up_src, l + 1, c_in)
                check_nn(self.lib.cm_nn_up_bwd(st, B, *d[l + 1], c_in, _vp(du), _vp(dnext)))
                dsrc = dnext
        for l, na, nb_, c_in, c in ((3, "e3a", "e3b", 64, 128), (2, "e2a", "e2b", 32, 64), (1, "e1a", "e1b", 16, 32)):
            dna = gz(na, l, c)
            self._conv_bwd(nb_, h[na], h["e%d" % l], de[l], d[l], B, dna)
            dp = gz("p%d" % l, l, c_in)
            self._conv_bwd(na, h["p%d" % l], h[na], dna, d[l], B, dp)
            check_nn(self.lib.cm_nn_pool_bwd(st, B, *d[l - 1], c_in, _vp(dp), _vp(de[l - 1])))
        dna = gz("e0a", 0, 16)
        self._conv_bwd("e0b", h["e0a"], h["e0"], de[0], d[0], B, dna)
        self._conv_bwd("e0a", h["f"], h["e0a"], dna, d[0], B, None)

    def zero_grad(self):
        for L in self.layers.values():
            L["gW"].zero_()
            L["gb"].zero_()

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step on every layer, then fresh BF16 compute copies."""
        self.step_count += 1
        for L in self.layers.values():
            for w, m, v, g in (("W", "mW", "vW", "gW"), ("b", "mb", "vb", "gb")):
                check_nn(self.lib.cm_nn_adam(self.stream, L[w].numel(), _vp(L[w]), _vp(L[m]), _vp(L[v]), _vp(L[g]),
                                             float(lr), b1, b2, eps, self.step_count))
        self.refresh()

    def loss(self, out, target, mask, M):
        """Masked MSE over the first 3 channels; returns (device FP64 partial sums, d loss / d out)."""
        dout = self._buf("dout", M * 16, torch.float32)
        nb = (M + 255) // 256
        part = self._buf("part", nb, torch.float64)
        count = float(mask[:M].sum().item()) if isinstance(mask, torch.Tensor) else float(np.sum(mask))
        check_nn(self.lib.cm_nn_mse(self.stream, M, OUT_CH, 16, _vp(out), _vp(target), _vp(mask),
                                    1.0 / max(1.0, 3.0 * count), _vp(dout), _vp(part)))
        return part, dout, count
