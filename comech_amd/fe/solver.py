"""Drop-in for the reference's ``fe/solver.py`` (voxel FE on a rectilinear grid), GPU path.

Same names, arguments and return values as ``fe/solver.py``:
``Grid``, ``lame``, ``ortho_coeffs``, ``element_matrices``, ``face_node_areas``, ``solve``,
``cell_stress``.  ``solve`` runs the matrix-free GMG-PCG of ``libcomech_hip.so``; there is no CPU
path.  ``U`` has NaN at nodes no active cell touches (fe/solver.py:248-249).

Differences, by design:
* ``method``: "gmg" (default here) and "cg" both run GMG-PCG on the GPU; "direct" solves with the
  dense inverse of the whole system, formed on the GPU (blocked Gauss-Jordan, FP64 matrix cores),
  for systems of at most 30,000 DOFs (larger ones raise ``ValueError``: use "gmg").
* ``info["K"]`` is a scipy ``LinearOperator`` backed by the GPU operator instead of an assembled
  CSR matrix (the GPU path never assembles).  ``info`` also carries the reference GPU solver's keys
  ``t_gpu`` (solve seconds), ``t_hier`` (multigrid set-up seconds) and ``coarse`` (coarsest-level
  method), as rb/gpusolve.py:259-260,285.
* ``coeffs`` (the z-orthotropic materials of fe/solver.py:43-74): the isotropic part (lambda =
  c1, mu = c4) runs in the matrix-free operator, and every active cell whose coefficients differ
  from ortho_coeffs(c1, c4) adds its exact element difference sum_k (c_k - c_k^iso) KB_k as extra
  stiffness (the parametric engine changes only c5 of cartilage cells,
  fe/simulate_parametric.py:197-201); stresses use D = sum_k c_k P_k.
* extra keyword options select the GPU solver: ``mode`` ("fp64", "ref" = FP64 PCG with FP32
  V-cycle, "mixed" = FP32 PCG with FP64 refinement), ``smoother`` ("chebyshev" as fe/gmg.py:61-76,
  or "jacobi"), ``omega``, ``npre``, ``npost``, ``levels``, ``min_dofs``, ``max_levels``,
  ``check_every``, ``inner_tol``.
"""
import numpy as np

from .. import engine as E
from ..grid import LOC as _LOC  # noqa: F401  (fe/solver.py:12 name)
from ..grid import Grid, corners_of, element_matrices, face_node_areas, lame, ortho_coeffs  # noqa: F401
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


DIRECT_MAX_DOFS = 30000


def ortho_split(grid, coeffs, active):
    """z-orthotropic coefficients (ncell, 6) -> (lam, mu, COO) with lam = c1, mu = c4 per cell and
    the COO of the element differences sum_k (c_k - c_k^iso(lam, mu)) KB_k (2x2x2 Gauss,
    fe/solver.py:63-74) of the active cells where they are non-zero, over node DOFs 3 node + comp."""
    c = np.asarray(coeffs, float)
    lam_i, mu_i = c[:, 1].copy(), c[:, 4].copy()
    d = c - ortho_coeffs(lam_i, mu_i)
    cells = np.flatnonzero(np.asarray(active, bool) & (np.abs(d).max(1) > 0))
    if len(cells) == 0:
        return lam_i, mu_i, None
    KB = grid.KB
    did = grid.dim_id[cells]
    Ke = np.empty((len(cells), 24, 24))
    for f in np.unique(did):
        m = did == f
        Ke[m] = np.einsum("nk,kab->nab", d[cells[m]], KB[f])
    dofs = (3 * corners_of(grid, cells)[:, :, None] + np.arange(3)).reshape(len(cells), 24)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    return lam_i, mu_i, (rows.astype(np.int64), cols.astype(np.int64), Ke.ravel())


def solve(grid, lam, mu, active, springs, loads, method="gmg", tol=1e-8, x0=None, maxiter=20000, penalties=(),
          coeffs=None, node_groups=None, U0=None, device=None, **opts):
    """fe/solver.py:202-254 on the GPU.  springs: [(node_ids, area, kvec[3])]; loads:
    [(node_ids, area, traction[3])]; penalties: [(nodes, weights, comp, k)]; coeffs: (ncell, 6)
    z-orthotropic coefficients (fe/solver.py:55-60) instead of lam/mu."""
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
    if coeffs is not None:
        lam, mu, ortho = ortho_split(grid, coeffs, active)
        parts.append(ortho)
    extra = E.cat_coo(parts)
    f = np.zeros(3 * len(nused))
    for nodes, area, t in loads:                                     # fe/solver.py:216-219
        nodes = np.asarray(nodes, np.int64)
        ln = nmap[nodes]
        ok = ln >= 0
        for c in range(3):
            np.add.at(f, 3 * ln[ok] + c, np.asarray(area, float)[ok] * t[c])
    if method == "direct":
        if 3 * len(nused) > DIRECT_MAX_DOFS:
            raise ValueError(f"method='direct' forms the dense inverse on the GPU: at most {DIRECT_MAX_DOFS} DOFs "
                             f"(this system has {3 * len(nused)}); use method='gmg'")
        mg = dict(mg, levels=1)
    P = E.Problem(grid, lam, mu, active, sri=False, extra=extra, node_groups=None if method == "direct" else node_groups,
                  device=device, **mg)
    if U0 is not None:
        x0 = np.nan_to_num(np.asarray(U0)[P.nused]).ravel()
    x0f = None if x0 is None else P.full_rhs(x0)
    if method == "direct":
        U, info, _ = P.direct(P.full_rhs(f))
    else:
        U, info, _ = P.solve(P.full_rhs(f), x0f, tol=tol, maxiter=maxiter, **so)
    info.update(ndof=P.ndof, K=E.matvec_operator(P), f=f, used=P.nused)
    return U, info


def cell_stress(grid, lam, mu, U, cells, coeffs=None, device=None):
    """fe/solver.py:257-272 on the GPU: (sig (n,6), vm (n,), eps (n,6)) at cell centres; with coeffs
    the z-orthotropic stress sig = sum_k c_k P_k eps (fe/solver.py:267-269)."""
    return cell_stress_gpu(grid, lam, mu, U, cells, device=device, coeffs=coeffs)
