"""Solve with the learned warm start.

The network predicts the displacement (comech_amd.learned.net, per cell, mapped to nodes).  A raw
prediction is not rigid inside the bones (17000 MPa beside 1 MPa soft tissue), so its energy-norm
error can exceed that of starting from zero.  The initial guess is therefore the Galerkin
projection of the solution onto a coarse space built from it: x0 = Z (Z^T A Z)^-1 Z^T b with
Z = [the prediction's x, y, z components on soft-tissue nodes, the six rigid modes of every bone]
(cm_nn_zt / cm_nn_zc, include/comech_learned.h) -- the energy-optimal combination, never worse than
zero in the A-norm -- the same form as the reference's additive coarse correction
Z E^-1 Z^T r (rb/gpusolve.py:217-228) with its rigid modes (fe/solver.py:177-199).  Plain GMG-PCG is
the fallback: unless the FP64 true residual has dropped 10x within 5 iterations
(cm_sys_set_fallback), the system is solved again from zero.  No reference counterpart (the
reference's rb engine ignores U0, fe/simulate.py:17).
"""
import ctypes
import time

import numpy as np
import torch

from .. import _lib
from .._lib import check_nn
from .data import encode, to_nodes

FLAG_FALLBACK = -2


def _vp(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def predict(net, enc):
    """The network's raw prediction as an FP64 node vector (the solver's level-0 layout)."""
    out = net.forward(enc["F"], enc["dims"], 1)
    return to_nodes(enc, out, float(net.meta["tscale"]))


class CoarseSpace:
    """Z of a system: node groups from the bone parts, group centres, and the prediction."""

    def __init__(self, enc, pred):
        self.lib = _lib.load()
        self.enc = enc
        inp = enc["inputs"]
        if "part" not in inp:
            raise ValueError("the description carries no bone parts")
        self.dev = pred.device
        self.st = ctypes.c_void_p(torch.cuda.current_stream(self.dev).cuda_stream)
        nx, ny, nz = enc["dims"]
        nn = enc["nn"]
        self.group = torch.empty(nn, dtype=torch.int32, device=self.dev)
        check_nn(self.lib.cm_nn_node_groups(self.st, nx, ny, nz, _vp(enc["lab"]), _vp(inp["part"]), _vp(enc["irr_of"]),
                                            _vp(enc["irr_nodes"]), nn, _vp(self.group)))
        self.ngroup = int(inp["part"].max().item())
        acc = torch.empty(4 * max(self.ngroup, 1), dtype=torch.float64, device=self.dev)
        xe, ye, ze = inp["edges"]
        self.args = (nx, ny, nz, nn, _vp(self.group), _vp(inp["copy_orig"]), _vp(xe), _vp(ye), _vp(ze))
        if self.ngroup:
            check_nn(self.lib.cm_nn_group_centres(self.st, *self.args, self.ngroup, _vp(acc)))
        a = acc.view(-1, 4)
        self.centre = (a[:, :3] / a[:, 3:].clamp(min=1.0)).contiguous()
        self.pred = pred
        self.k = 3 + 6 * self.ngroup

    def zt(self, v):
        out = torch.zeros(self.k, dtype=torch.float64, device=self.dev)
        check_nn(self.lib.cm_nn_zt(self.st, *self.args, self.ngroup, _vp(self.centre), _vp(self.pred), _vp(v), _vp(out)))
        return out

    def zc(self, c, x=None):
        x = torch.empty_like(self.pred) if x is None else x
        check_nn(self.lib.cm_nn_zc(self.st, *self.args, self.ngroup, _vp(self.centre), _vp(self.pred), _vp(c), _vp(x)))
        return x

    def galerkin(self, sys_, b):
        """x0 = Z (Z^T A Z)^+ Z^T b (E formed with k products of the system's FP64 operator)."""
        E = torch.empty((self.k, self.k), dtype=torch.float64, device=self.dev)
        z = torch.empty_like(self.pred)
        az = torch.empty_like(self.pred)
        e = torch.zeros(self.k, dtype=torch.float64, device=self.dev)
        for j in range(self.k):
            e.zero_()
            e[j] = 1.0
            self.zc(e, z)
            sys_.apply(z, az)
            E[:, j] = self.zt(az)
        f = self.zt(b)
        Eh = E.cpu().numpy()
        c = np.linalg.pinv(0.5 * (Eh + Eh.T)) @ f.cpu().numpy()
        return self.zc(torch.as_tensor(c, device=self.dev))


def learned_solve(sys_, desc, b, net, tol=1e-8, mode="ref", maxiter=3000, project=True, fallback_iters=5, drop=10.0,
                  enc=None):
    """Solve sys_ x = b from the learned initial guess (projected onto its coarse space unless
    project=False; net=None: the same projection without the network, i.e. onto the bones' rigid
    modes alone).  Returns (x, info) with info["fallback"] (True when the early-exit rule handed
    over to plain GMG-PCG), info["iters"] (all iterations of both attempts), info["t_guess_ms"]
    (prediction and projection, wall) and info["r0_rel"] (true residual of the guess)."""
    t0 = time.perf_counter()
    enc = encode(desc, sys_.device) if enc is None else enc
    x0 = predict(net, enc) if net is not None else torch.zeros(3 * enc["nn"], dtype=torch.float64, device=sys_.device)
    if project:
        x0 = CoarseSpace(enc, x0).galerkin(sys_, b)
    r0 = float(torch.linalg.vector_norm(b - sys_.apply(x0)) / torch.linalg.vector_norm(b))
    t_guess = (time.perf_counter() - t0) * 1e3
    sys_.set_fallback(fallback_iters, drop)
    try:
        x, info = sys_.solve(b, x0, tol=tol, maxiter=maxiter, mode=mode)
    finally:
        sys_.set_fallback(0)
    info.update(fallback=False, t_guess_ms=t_guess, r0_rel=r0)
    if info["flag"] == FLAG_FALLBACK:
        first = info["iters"]
        x, info = sys_.solve(b, None, tol=tol, maxiter=maxiter, mode=mode)
        info.update(iters=info["iters"] + first, fallback=True, t_guess_ms=t_guess, r0_rel=r0)
    return x, info
