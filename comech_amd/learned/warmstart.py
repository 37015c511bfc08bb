"""Solve with the learned warm start: the network's prediction is the initial guess of PCG and,
optionally, spans an additive coarse correction z += Z E^-1 Z^T r (the form of
rb/gpusolve.py:217-228) with Z = the prediction's three displacement components (E = Z^T A Z, so
the preconditioner stays symmetric positive definite).  Plain GMG-PCG is the fallback: unless the
FP64 true residual has dropped 10x within 5 iterations (cm_sys_set_fallback), the system is solved
again from zero without the prediction.  No reference counterpart (the reference's rb engine
ignores U0, fe/simulate.py:17)."""
import time

import numpy as np
import torch

from .data import encode, to_nodes

FLAG_FALLBACK = -2


def predict(net, enc):
    """The network's initial guess as an FP64 node vector (the solver's level-0 layout)."""
    out = net.forward(enc["F"], enc["dims"], 1)
    return to_nodes(enc, out, float(net.meta["tscale"]))


def coarse_space(x0):
    """Z as COO over node DOFs: column c holds component c of the prediction."""
    v = x0.cpu().numpy()
    rows = np.flatnonzero(v)
    return rows, rows % 3, v[rows], 3


def learned_solve(sys_, desc, b, net, tol=1e-8, mode="ref", maxiter=3000, correction=False, fallback_iters=5,
                  drop=10.0, enc=None):
    """Solve sys_ x = b from the network's prediction.  Returns (x, info) with info["fallback"]
    (True when the early-exit rule handed over to plain GMG-PCG), info["iters"] (all iterations,
    both attempts), info["t_predict_ms"] and info["r0_rel"] (true residual of the prediction)."""
    t0 = time.perf_counter()
    enc = encode(desc, sys_.device) if enc is None else enc
    x0 = predict(net, enc)
    r0 = float(torch.linalg.vector_norm(b - sys_.apply(x0)) / torch.linalg.vector_norm(b))
    if correction:
        sys_.set_rigid(*coarse_space(x0))
    t_pred = (time.perf_counter() - t0) * 1e3
    sys_.set_fallback(fallback_iters, drop)
    try:
        x, info = sys_.solve(b, x0, tol=tol, maxiter=maxiter, mode=mode)
    finally:
        sys_.set_fallback(0)
    info.update(fallback=False, t_predict_ms=t_pred, r0_rel=r0)
    if info["flag"] == FLAG_FALLBACK:
        first = info["iters"]
        if correction:
            sys_.clear_rigid()
        x, info = sys_.solve(b, None, tol=tol, maxiter=maxiter, mode=mode)
        info.update(iters=info["iters"] + first, fallback=True, t_predict_ms=t_pred, r0_rel=r0)
    elif correction:
        sys_.clear_rigid()
    return x, info
