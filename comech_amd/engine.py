"""Host logic shared by the drop-in solvers: reference inputs -> GPU system -> reference outputs.

The reference assembles a CSR matrix on compact DOFs ``3*nmap[node]+c`` (fe/solver.py:137-162,
rb/splitfe.py:72-91) and hands it to ``pcg_gmg``.  Here nothing is assembled: the cells stay as
labels/materials on the GPU, and only the genuinely sparse terms (springs, penalties, ligament and
interface springs, the ``extra(nmap, n)`` callback) become a small COO over level-0 node DOFs
``3*node+c``.  Node ids are the reference's own (grid ids, then split copies), so the level-0
vector *is* the reference's ``U`` array (nnode_total, 3) without the NaN of unused nodes.
"""
import time

import numpy as np
import scipy.sparse as sp
import torch

from .grid import corners_of
from .split import irregular_cells
from .system import VoxelSystem, cell_stress_gpu, require_gpu


def split_cells(grid, active, cn):
    """Active cells whose connectivity differs from the grid corners (split-copy cells).
    Returns (irr_cells, irr_nodes(n,8))."""
    return irregular_cells(grid, active, cn)


def used_nodes(grid, active, nn, irr_cells=None, irr_nodes=None):
    """Mask of nodes touched by active cells (fe/solver.py:141, rb/splitfe.py:73)."""
    used = np.zeros(nn, bool)
    cells = np.flatnonzero(active)
    if irr_cells is not None and len(irr_cells):
        keep = np.ones(grid.ncell, bool)
        keep[irr_cells] = False
        cells = cells[keep[cells]]
        used[np.asarray(irr_nodes).ravel()] = True
    for a in range(0, len(cells), 1 << 20):
        used[corners_of(grid, cells[a:a + (1 << 20)]).ravel()] = True
    return used


def compact_to_full_coo(M, nused):
    """CSR/COO on compact DOFs -> COO on node DOFs 3*node+c."""
    M = sp.coo_matrix(M)
    r = 3 * nused[M.row // 3] + M.row % 3
    c = 3 * nused[M.col // 3] + M.col % 3
    return r.astype(np.int64), c.astype(np.int64), M.data.astype(np.float64)


def diag_coo(nodes, kvals, comps=(0, 1, 2)):
    """Diagonal springs: nodes (m,), kvals (m, len(comps)) -> COO."""
    rows = (3 * np.asarray(nodes, np.int64)[:, None] + np.asarray(comps)[None, :]).ravel()
    return rows, rows.copy(), np.asarray(kvals, float).ravel()


def rank1_coo(dofs, w, k):
    """k w w^T on the given DOFs (duplicates in dofs are summed first)."""
    dofs = np.asarray(dofs, np.int64)
    w = np.asarray(w, float)
    ud, inv = np.unique(dofs, return_inverse=True)
    wu = np.zeros(len(ud))
    np.add.at(wu, inv, w)
    nz = wu != 0
    ud, wu = ud[nz], wu[nz]
    return np.repeat(ud, len(ud)), np.tile(ud, len(ud)), (k * np.outer(wu, wu)).ravel()


def cat_coo(parts):
    parts = [p for p in parts if p is not None and len(p[0])]
    if not parts:
        return None
    return tuple(np.concatenate([p[q] for p in parts]) for q in range(3))


def rigid_coo(xyz, used, node_groups):
    """Rigid-body modes per node group (fe/solver.py:177-199, rb/splitfe.py:105-120) as COO over
    node DOFs: 3 translations and 3 rotations with r/20, groups of fewer than 4 used nodes skipped."""
    rows, cols, vals = [], [], []
    col = 0
    for nodes in node_groups:
        nodes = np.asarray(nodes, np.int64)
        nodes = nodes[used[nodes]]
        if len(nodes) < 4:
            continue
        r = (xyz[nodes] - xyz[nodes].mean(0)) / 20.0
        one = np.ones(len(nodes))
        modes = [((0, one),), ((1, one),), ((2, one),),
                 ((1, -r[:, 2]), (2, r[:, 1])), ((0, r[:, 2]), (2, -r[:, 0])), ((0, -r[:, 1]), (1, r[:, 0]))]
        for m in modes:
            for comp, v in m:
                rows.append(3 * nodes + comp); cols.append(np.full(len(nodes), col)); vals.append(v)
            col += 1
    if col == 0:
        return None, 0
    return (np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)), col


def fragment_groups(grid, part, active, irr_cells=(), irr_nodes=None, min_cells=8):
    """Node groups of the fragments of split bones: for every bone part (part > 0, reference cell
    order) whose active cells form two or more face-connected pieces (an osteotomy cut, a resected
    and filled gap), the nodes of each piece of at least ``min_cells`` cells -- the split copies of
    irregular cells included (irr_nodes, 8 per listed cell).  Rigid modes of these groups
    (rigid_coo) as the additive coarse correction of the preconditioner let the fragments move
    apart on every level, which the grid hierarchy cannot represent (no reference counterpart: the
    reference's engine takes one group per bone, rb/rbfe.py:180)."""
    import scipy.ndimage as ndi
    from .grid import corners_of
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    part = np.asarray(part).reshape(nx, ny, nz)
    act = np.asarray(active, bool).reshape(nx, ny, nz)
    pm = np.where(act, part, 0).astype(np.int32)
    irr_cells = np.asarray(irr_cells, np.int64)
    iorder = np.argsort(irr_cells, kind="stable")
    isorted = irr_cells[iorder]
    out = []
    for p, sl in enumerate(ndi.find_objects(pm), start=1):
        if sl is None:
            continue
        comp, n = ndi.label(pm[sl] == p)     # inside the part's bounding box
        if n < 2:
            continue
        base = np.array([s_.start for s_ in sl])
        for q in range(1, n + 1):
            idx = np.argwhere(comp == q)
            if len(idx) < min_cells:
                continue
            ijk = idx + base
            cells = (ijk[:, 0] * ny + ijk[:, 1]) * nz + ijk[:, 2]
            nodes = corners_of(grid, cells)
            if len(isorted):
                pos = np.minimum(np.searchsorted(isorted, cells), len(isorted) - 1)
                hit = isorted[pos] == cells
                nodes[hit] = np.asarray(irr_nodes)[iorder[pos[hit]]]
            out.append(np.unique(nodes.ravel()))
    return out


# coarsest-level method (the "coarse" key of rb/gpusolve.py:259-260): a dense FP64 inverse formed on
# the GPU by blocked Gauss-Jordan, applied as a matrix-vector product
COARSE = "dense-inverse"

MG_KEYS = ("min_dofs", "max_levels", "levels", "smoother", "degree", "omega", "npre", "npost", "copy_blocks")
SOLVE_KEYS = ("mode", "check_every", "inner_tol")


def split_opts(kw):
    mg = {k: kw.pop(k) for k in list(kw) if k in MG_KEYS}
    so = {k: kw.pop(k) for k in list(kw) if k in SOLVE_KEYS}
    if kw:
        raise TypeError(f"unexpected options: {sorted(kw)}")
    return mg, so


class Problem:
    """One voxel FE problem on the GPU in reference terms."""

    def __init__(self, grid, lam, mu, active, *, sri, cn=None, orig=None, extra=None, rank1=None, node_groups=None,
                 device=None, **mg):
        self.grid = grid
        self.lam = np.asarray(lam, float)
        self.mu = np.asarray(mu, float)
        self.active = np.asarray(active, bool)
        self.sri = sri
        self.cn = cn
        if cn is not None:
            orig = np.asarray(orig, np.int64)
            self.copy_orig = orig[grid.nnode:]
            self.irr_cells, self.irr_nodes = split_cells(grid, self.active, cn)
        else:
            self.copy_orig = np.zeros(0, np.int64)
            self.irr_cells, self.irr_nodes = np.zeros(0, np.int64), np.zeros((0, 8), np.int64)
        self.nn = grid.nnode + len(self.copy_orig)
        self.used = used_nodes(grid, self.active, self.nn, self.irr_cells, self.irr_nodes)
        self.nused = np.flatnonzero(self.used)
        self.nmap = -np.ones(self.nn, np.int64)
        self.nmap[self.nused] = np.arange(len(self.nused))
        self.ndof = 3 * len(self.nused)
        t0 = time.perf_counter()
        self.sys = VoxelSystem(grid, sri=sri, active=self.active, lam=self.lam, mu=self.mu,
                               copy_orig=self.copy_orig, irr_cells=self.irr_cells, irr_nodes=self.irr_nodes,
                               extra=extra, rank1=rank1, device=device, **mg)
        self.t_hier = time.perf_counter() - t0
        self.device = self.sys.device
        if node_groups:
            xyz = grid.node_xyz if cn is None else grid.node_xyz[np.asarray(orig)]
            Z, ncol = rigid_coo(xyz, self.used, node_groups)
            if ncol:
                self.sys.set_rigid(*Z, ncol)

    def full_rhs(self, f_compact):
        b = np.zeros(3 * self.nn)
        b.reshape(-1, 3)[self.nused] = np.asarray(f_compact).reshape(-1, 3)
        return torch.from_numpy(b).to(self.device)

    def solve(self, b_full, x0_full=None, tol=1e-8, maxiter=800, **so):
        x, info = self.sys.solve(b_full, x0_full, tol=tol, maxiter=maxiter, **so)
        info.update(t_gpu=round(info["t_solve_ms"] / 1e3, 4), t_hier=round(self.t_hier, 4), coarse=COARSE)
        return self._host(x), info, x

    def direct(self, b_full):
        """x = A^-1 b with the dense inverse of a one-level system (the coarsest-level solve of the
        V-cycle), and its FP64 true residual."""
        if len(self.sys.levels()) != 1:
            raise ValueError("direct needs a one-level system (levels=1)")
        t0 = time.perf_counter()
        x = self.sys.vcycle(b_full)
        r = float(torch.linalg.vector_norm(b_full - self.sys.apply(x)) /
                  max(float(torch.linalg.vector_norm(b_full)), 1e-300))
        torch.cuda.synchronize(self.device)
        info = dict(iters=0, flag=0, outer=0, levels=1, relres=r, mode="direct", device=str(self.device),
                    t_gpu=round(time.perf_counter() - t0, 4), t_hier=round(self.t_hier, 4), coarse=COARSE)
        return self._host(x), info, x

    def _host(self, x):
        U = x.view(-1, 3).cpu().numpy().copy()
        U[~self.used] = np.nan
        return U

    def close(self):
        self.sys.close()


def matvec_operator(problem):
    """scipy LinearOperator on compact DOFs backed by the GPU operator (used for info['K'])."""
    import scipy.sparse.linalg as spla
    P = problem

    def mv(v):
        full = np.zeros(3 * P.nn)
        full.reshape(-1, 3)[P.nused] = np.asarray(v).reshape(-1, 3)
        y = P.sys.apply(torch.from_numpy(full).to(P.device)).view(-1, 3).cpu().numpy()
        return y[P.nused].ravel()
    return spla.LinearOperator((P.ndof, P.ndof), matvec=mv, dtype=np.float64)
