"""Drop-in for the reference's ``fe/solver.py`` (voxel FE on a rectilinear grid), GPU path.

Same names, arguments and return values as ``fe/solver.py``:
``Grid``, ``lame``, ``ortho_coeffs``, ``element_matrices``, ``face_node_areas``, ``solve``,
``cell_stress``.  ``solve`` runs the matrix-free GMG-PCG of ``libcomech_hip.so``; there is no CPU
path.  ``U`` has NaN at nodes no active cell touches (fe/solver.py:248-249).

Differences, by design:
* ``method``: "gmg" (default here) and "cg" both run GMG-PCG on the GPU; "direct" runs GMG-PCG to
  a relative residual of 1e-12 (there is no sparse direct factorisation on the GPU path).
* ``info["K"]`` is a scipy ``LinearOperator`` backed by the GPU operator instead of an assembled
  CSR matrix (the GPU path never assembles).
* extra keyword options select the GPU solver: ``mode`` ("fp64", "ref" = FP64 PCG with FP32
  V-cycle, "mixed" = FP32 PCG with FP64 refinement), ``smoother`` ("chebyshev" as fe/gmg.py:61-76,
  or "jacobi"), ``omega``, ``npre``, ``npost``, ``levels``, ``min_dofs``, ``max_levels``,
  ``check_every``, ``inner_tol``.
* ``coeffs`` (the z-orthotropic basis of the parametric engine) is not supported on the GPU path
  and raises ``NotImplementedError``.
"""
import numpy as np

from .. import engine as E
from ..grid import LOC as _LOC  # noqa: F401  (fe/solver.py:12 name)
from ..grid import Grid, element_matrices, face_node_areas, lame, ortho_coeffs  # noqa: F401
from ..system import cell_stress_gpu


def rigid_modes(grid, nmap, n, node_groups):
    """Sparse Z (n x 6*groups) on compact DOFs (fe/solver.py:177-199)."""
    import scipy.sparse as sp
    used = nmap >= 0
    Z, ncol = E.rigid_coo(grid.node_xyz, used, node_groups)
    if not ncol:
        return sp.csr_matrix((n, 0))
    r, c, v = Z
    return sp.csr_matrix((v, (3 * nmap[r // 3] + r % 3, c)), shape=(n, ncol))


def solve(grid, lam, mu, active, springs, loads, method="gmg", tol=1e-8, x0=None, maxiter=20000, penalties=(),
          coeffs=None, node_groups=None, U0=None, device=None, **opts):
    """fe/solver.py:202-254 on the GPU.  springs: [(node_ids, area, kvec[3])]; loads:
    [(node_ids, area, traction[3])]; penalties: [(nodes, weights, comp, k)]."""
    if coeffs is not None:
        raise NotImplementedError("orthotropic coeffs are not supported by the GPU solver")
    if method not in ("gmg", "cg", "direct"):
        raise ValueError(f"unknown method {method!r}")
    mg, so = E.split_opts(dict(opts))
    so.setdefault("mode", "fp64")
    so.setdefault("check_every", 1)
    active = np.asarray(active, bool)
    nn = grid.nnode
    used = E.used_nodes(grid, active, nn)
    nmap = -np.ones(nn, np.int64)
    nused = np.flatnonzero(used)
    nmap[nused] = np.arange(len(nused))
    parts = []
    for nodes, area, kvec in springs:                               # fe/solver.py:212-215
        nodes = np.asarray(nodes, np.int64)
        ok = used[nodes]
        parts.append(E.diag_coo(nodes[ok], np.asarray(area, float)[ok][:, None] * np.asarray(kvec, float)[None, :]))
    for nodes, weights, comp, k in penalties:                        # fe/solver.py:224-232
        nodes = np.asarray(nodes, np.int64)
        ok = used[nodes]
        w = np.asarray(weights, float)[ok]
        dofs = 3 * nodes[ok] + comp
        ud, inv = np.unique(dofs, return_inverse=True)
        wu = np.zeros(len(ud)); np.add.at(wu, inv, w)
        wu /= wu.sum()
        parts.append(E.rank1_coo(ud, wu, k))
    extra = E.cat_coo(parts)
    f = np.zeros(3 * len(nused))
    for nodes, area, t in loads:                                     # fe/solver.py:216-219
        nodes = np.asarray(nodes, np.int64)
        ln = nmap[nodes]
        ok = ln >= 0
        for c in range(3):
            np.add.at(f, 3 * ln[ok] + c, np.asarray(area, float)[ok] * t[c])
    P = E.Problem(grid, lam, mu, active, sri=False, extra=extra, node_groups=node_groups, device=device, **mg)
    if U0 is not None:
        x0 = np.nan_to_num(np.asarray(U0)[P.nused]).ravel()
    x0f = None if x0 is None else P.full_rhs(x0)
    if method == "direct":
        tol = min(tol, 1e-12)
    U, info, _ = P.solve(P.full_rhs(f), x0f, tol=tol, maxiter=maxiter, **so)
    info.update(ndof=P.ndof, K=E.matvec_operator(P), f=f, used=P.nused)
    return U, info


def cell_stress(grid, lam, mu, U, cells, coeffs=None, device=None):
    """fe/solver.py:257-272 on the GPU: (sig (n,6), vm (n,), eps (n,6)) at cell centres."""
    if coeffs is not None:
        raise NotImplementedError("orthotropic coeffs are not supported by the GPU solver")
    return cell_stress_gpu(grid, lam, mu, U, cells, device=device)
