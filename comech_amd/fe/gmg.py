"""Drop-in for the reference's ``fe/gmg.py``: ``pcg_gmg`` on an assembled CSR matrix, on the GPU.

``pcg_gmg(A, b, used_nodes, grid_edges, tol, maxiter, x0, A_op, Z)`` keeps the reference signature
and result ``(x, {iters, flag, levels})`` (fe/gmg.py:120-137).  The CSR matrix becomes the sparse
"extra" operator of a GPU system on the reference grid: compact DOF ``3*i+c`` lives at node
``used_nodes[i]``; a grid node listed more than once (split copies, whose ``used_nodes`` entry is
the original grid node, rb/splitfe.py:153) gets one copy node per repeat, interpolated like its
original on coarse levels.  Coarse operators are the Galerkin products ``P^T A P`` with the
trilinear ``P`` of fe/gmg.py:7-26, formed on the host once per system as in rb/gpusolve.py:39-65;
smoothing, V-cycle, coarse solve and PCG run on the GPU.  Options as ``comech_amd.fe.solver.solve``
(default: FP64 throughout, residual checked every iteration as scipy's ``cg``).
"""
import time

import numpy as np
import scipy.sparse as sp
import torch

from .. import engine as E
from ..grid import Grid
from ..system import VoxelSystem


def csr_system(A, used_nodes, grid_edges, device=None, **mg):
    g = Grid(*grid_edges)
    used_nodes = np.asarray(used_nodes, np.int64)
    order = np.argsort(used_nodes, kind="stable")
    first = np.ones(len(used_nodes), bool)
    srt = used_nodes[order]
    first[order[1:]] = srt[1:] != srt[:-1]
    node_of = np.empty(len(used_nodes), np.int64)
    node_of[first] = used_nodes[first]
    rep = np.flatnonzero(~first)
    node_of[rep] = g.nnode + np.arange(len(rep))
    copy_orig = np.concatenate([np.arange(g.nnode), used_nodes[rep]])
    extra = E.compact_to_full_coo(A, node_of)
    nn = g.nnode + len(rep)
    used = np.zeros(nn, np.uint8)
    used[node_of] = 1
    sys_ = VoxelSystem(g, sri=True, active=np.zeros(g.ncell, bool), labels=np.full(g.ncell, 255, np.uint8),
                       lam_tab=np.zeros(1), mu_tab=np.zeros(1), copy_orig=copy_orig[g.nnode:], extra=extra,
                       used=used, device=device, **mg)
    return sys_, node_of


def pcg_gmg(A, b, used_nodes, grid_edges, tol=1e-7, maxiter=500, x0=None, A_op=None, Z=None, device=None, **opts):
    """fe/gmg.py:120-137 on the GPU.  Returns (x, dict(iters, flag, levels, ...))."""
    if A_op is not None:
        raise NotImplementedError("A_op (a separate operator for CG) is not supported by the GPU solver")
    mg, so = E.split_opts(dict(opts))
    so.setdefault("mode", "fp64")
    so.setdefault("check_every", 1)
    A = sp.csr_matrix(A)
    t0 = time.perf_counter()
    sys_, node_of = csr_system(A, used_nodes, grid_edges, device=device, **mg)
    t_hier = time.perf_counter() - t0
    try:
        if Z is not None:
            Zc = sp.coo_matrix(Z)
            sys_.set_rigid(3 * node_of[Zc.row // 3] + Zc.row % 3, Zc.col, Zc.data, Zc.shape[1])
        n = 3 * sys_.nn
        dofs = (3 * node_of[:, None] + np.arange(3)).ravel()
        bf = np.zeros(n); bf[dofs] = np.asarray(b, float)
        xf = None
        if x0 is not None:
            xf = np.zeros(n); xf[dofs] = np.asarray(x0, float)
            xf = torch.from_numpy(xf).to(sys_.device)
        x, info = sys_.solve(torch.from_numpy(bf).to(sys_.device), xf, tol=tol, maxiter=maxiter, **so)
        info.update(t_gpu=round(info["t_solve_ms"] / 1e3, 4), t_hier=round(t_hier, 4), coarse=E.COARSE)
        return x.cpu().numpy()[dofs], info
    finally:
        sys_.close()
