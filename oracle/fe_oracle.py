"""FP64 CPU oracle of the CoMech voxel FE solve path (numpy / scipy restatement).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  The product never imports this module.

How each part is pinned (checked in ``tests/test_oracle.py``), so that a dropped term, a wrong sign,
a wrong index or a transposed operand fails at least one check:

==========================  =====================================================================
function                    pinned by
==========================  =====================================================================
Grid numbering              closed forms of fe/solver.py:104-109,128 on hand-checked small grids
element_matrices            symmetry; 6 rigid zero modes; energy of a linear field = V * e:D:e
                            (closed form); rank(KL)=rank of full-integration lambda part
kl_sri                      rank 1; equals V * (div at centre)^2 closed form
assemble / assemble_cn      brute-force element loop summation on tiny grids
split_connectivity(_fast)   the vectorised form against the loop form (rb/splitfe.py:262 says
                            "same result") on random rasters, plus hand-built 2-bone cases
p1d                         reproduces linear functions exactly; partition of unity
prolongation(axes)          all axes kept = identity; x kept = reference prolongation of y, z
                            kron identity; linear fields reproduced
Level.smooth (Chebyshev)    textbook Chebyshev error polynomial p(MA) by eigendecomposition
vcycle                      two-grid error-propagation formula on small dense systems
pcg_gmg / solve             scipy ``splu`` direct solve (library routine) and true residual
patch test                  fe/test_patch.py:4-14 closed forms uz_top=-10/E, ux=nu*10/E, vm=1
cell_stress(_cn)            uniaxial / hydrostatic / pure shear closed forms for von Mises
spring_matrix_cn            k w w^T closed form on hand-built springs
sample_fe                   constant and linear fields reproduced exactly
sample_nodes                nodes in active cells: linear fields reproduced exactly; nodes in
                            empty cells: the cell used is the brute-force nearest active cell
                            (smallest squared index distance, ties to the smallest (k, j, i),
                            the rule scipy's feature transform follows) and the value is that
                            cell's trilinear field at the clamped local coordinates
lattice_materials           single-phase cells return that phase's (E, nu); a hard phase in
                            series with one layer is the Reuss average of constrained and shear
                            moduli; two phases in a non-joint cell the Voigt average; joint cells
                            against a brute-force 3x3x3 neighbourhood scan; activity thresholds
bone_groups / bone_share    hand-built label maps; shares sum to the bone fraction
lattice_part                closed form on hand-built shares
bone_node_sets              hand-built part maps (node sets by enumeration)
ground_contact / leg_loads  contact: areas sum to the contact-weighted bottom face area; loads:
                            the resultant is -Fz along z and its moment puts it on the load line
                            (closed form), uniform when the line passes through the centroid
spring_end_bones            scipy cKDTree nearest-neighbour distances (library routine)
rigid_modes                 K Z = 0 for the 6 modes of a spring-free body; per-bone blocks
element_forces_cn           their scatter-sum equals the assembled K u
solve(penalties)            each penalty block is k w w^T with sum(w) = 1 (closed form)
coeffs (orthotropic)        energy of an affine field = V eps^T D eps; centre stress D eps;
                            isotropic coefficients reproduce lam KL + mu KM; column compliance
region_labels               hand-built cells covering every rule and its precedence
bone_forces                 uniform stress across a face: force = sigma n A (closed form)
ground_zones                hand-built reactions and zone edges
rigid_fit                   exact recovery of t, w from a rigid displacement field
effects                     hand-built fields: p99 / weighted means by the mask rules, shares,
                            rigid motion of the talus relative to the calcaneus recovered
==========================  =====================================================================
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ---------------------------------------------------------------------------------------------
# fe/solver.py:11-12  local corner order l = a + 2b + 4c
LOC = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=np.int64)


def shape_grad(xi, eta, zeta, dims):
    """Physical gradients (8,3) of the trilinear shape functions at a natural point
    (fe/solver.py:15-23)."""
    s = 2.0 * LOC - 1.0
    nat = np.array([xi, eta, zeta], float)
    out = np.empty((8, 3))
    for d in range(3):
        o1, o2 = [k for k in range(3) if k != d]
        out[:, d] = s[:, d] * (1 + s[:, o1] * nat[o1]) * (1 + s[:, o2] * nat[o2]) / 8.0 * (2.0 / dims[d])
    return out


def bmat(dN):
    """6x24 strain-displacement matrix, engineering shear order xx,yy,zz,xy,yz,xz
    (fe/solver.py:26-36)."""
    B = np.zeros((6, 24))
    for n in range(8):
        gx, gy, gz = dN[n]
        c = 3 * n
        B[0, c] = gx
        B[1, c + 1] = gy
        B[2, c + 2] = gz
        B[3, c], B[3, c + 1] = gy, gx
        B[4, c + 1], B[4, c + 2] = gz, gy
        B[5, c], B[5, c + 2] = gz, gx
    return B


D_LAMBDA = np.zeros((6, 6)); D_LAMBDA[:3, :3] = 1.0          # fe/solver.py:39
D_MU = np.diag([2.0, 2.0, 2.0, 1.0, 1.0, 1.0])               # fe/solver.py:40


def ortho_patterns():
    """z-orthotropic constitutive basis P_k (fe/solver.py:43-52)."""
    P = np.zeros((6, 6, 6))
    P[0, 0, 0] = P[0, 1, 1] = 1
    P[1, 0, 1] = P[1, 1, 0] = 1
    for (i, j) in ((0, 2), (2, 0), (1, 2), (2, 1)):
        P[2, i, j] = 1
    P[3, 2, 2] = 1
    P[4, 3, 3] = 1
    P[5, 4, 4] = P[5, 5, 5] = 1
    return P


def ortho_coeffs(lam, mu, shear_factor=None):
    """fe/solver.py:55-60."""
    c = np.stack([lam + 2 * mu, lam, lam, lam + 2 * mu, mu, mu], 1)
    if shear_factor is not None:
        c[:, 5] *= shear_factor
    return c


_GP = 1.0 / np.sqrt(3.0)


def _gauss_points():
    return [(x, y, z) for x in (-_GP, _GP) for y in (-_GP, _GP) for z in (-_GP, _GP)]


def element_matrices(dims):
    """(K_lambda, K_mu, B_centre) with 2x2x2 Gauss (fe/solver.py:77-90)."""
    dims = np.asarray(dims, float)
    w = dims.prod() / 8.0
    KL = np.zeros((24, 24)); KM = np.zeros((24, 24))
    for (x, y, z) in _gauss_points():
        B = bmat(shape_grad(x, y, z, dims))
        KL += w * B.T @ D_LAMBDA @ B
        KM += w * B.T @ D_MU @ B
    return KL, KM, bmat(shape_grad(0.0, 0.0, 0.0, dims))


def element_basis(dims):
    """Orthotropic element basis (6,24,24) (fe/solver.py:63-74)."""
    dims = np.asarray(dims, float)
    w = dims.prod() / 8.0
    P = ortho_patterns()
    K = np.zeros((6, 24, 24))
    for (x, y, z) in _gauss_points():
        B = bmat(shape_grad(x, y, z, dims))
        for k in range(6):
            K[k] += w * B.T @ P[k] @ B
    return K


def kl_sri(dims):
    """One-point (cell centre) lambda matrix, rank 1 (rb/splitfe.py:58-69)."""
    dims = np.asarray(dims, float)
    Bc = bmat(shape_grad(0.0, 0.0, 0.0, dims))
    return dims.prod() * Bc.T @ D_LAMBDA @ Bc


class Grid:
    """Rectilinear grid (fe/solver.py:93-128): nodes x-fastest, cells z-fastest."""

    def __init__(self, xe, ye, ze):
        self.xe, self.ye, self.ze = (np.asarray(e, float) for e in (xe, ye, ze))
        self.nx, self.ny, self.nz = len(self.xe) - 1, len(self.ye) - 1, len(self.ze) - 1
        self.shape = (self.nx, self.ny, self.nz)
        self.ncell = self.nx * self.ny * self.nz
        self.nnode = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.dx, self.dy, self.dz = np.diff(self.xe), np.diff(self.ye), np.diff(self.ze)
        # cell (i,j,k) -> id (i*ny + j)*nz + k  (fe/solver.py:104-105,128)
        cid = np.arange(self.ncell)
        self.cK = cid % self.nz
        self.cJ = (cid // self.nz) % self.ny
        self.cI = cid // (self.ny * self.nz)
        # node (i,j,k) -> i + (nx+1)(j + (ny+1)k)  (fe/solver.py:107)
        self.cell_nodes = np.stack([self.nid(self.cI + a, self.cJ + b, self.cK + c) for a, b, c in LOC], 1)
        dims = np.stack([self.dx[self.cI], self.dy[self.cJ], self.dz[self.cK]], 1)
        self.udims, inv = np.unique(np.round(dims, 6), axis=0, return_inverse=True)   # fe/solver.py:112
        self.dim_id = inv.ravel()
        mats = [element_matrices(d) for d in self.udims]
        self.KL = np.stack([m[0] for m in mats])
        self.KM = np.stack([m[1] for m in mats])
        self.Bc = np.stack([m[2] for m in mats])
        self.cell_vol = dims.prod(1)
        nid = np.arange(self.nnode)
        ni = nid % (self.nx + 1)
        nj = (nid // (self.nx + 1)) % (self.ny + 1)
        nk = nid // ((self.nx + 1) * (self.ny + 1))
        self.node_xyz = np.stack([self.xe[ni], self.ye[nj], self.ze[nk]], 1)
        self.cell_xyz = np.stack([(self.xe[:-1] + self.xe[1:])[self.cI] / 2, (self.ye[:-1] + self.ye[1:])[self.cJ] / 2,
                                  (self.ze[:-1] + self.ze[1:])[self.cK] / 2], 1)

    def nid(self, i, j, k):
        return i + (self.nx + 1) * (j + (self.ny + 1) * k)

    def cid(self, i, j, k):
        return (i * self.ny + j) * self.nz + k

    @property
    def KL_sri(self):
        return np.stack([kl_sri(d) for d in self.udims])

    @property
    def KB(self):
        return np.stack([element_basis(d) for d in self.udims])


def lame(E, nu):
    """fe/solver.py:131-134."""
    E = np.asarray(E, float); nu = np.asarray(nu, float)
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def _coo_to_csr(Ke, dofs, n):
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    return sp.coo_matrix((Ke.ravel(), (rows, cols)), shape=(n, n)).tocsr()


def assemble(grid, lam, mu, active, coeffs=None, KLf=None, cn=None):
    """Global CSR stiffness on compact DOFs 3*nmap[node]+c (fe/solver.py:137-162 and, with
    ``cn``/``KLf``, rb/splitfe.py:72-91).  Returns K, nmap, used, cells, dofs."""
    cells = np.nonzero(active)[0]
    conn = grid.cell_nodes if cn is None else cn
    c8 = conn[cells]
    used = np.unique(c8)
    nmap = -np.ones(max(grid.nnode, int(conn.max()) + 1), np.int64)
    nmap[used] = np.arange(len(used))
    dofs = (3 * nmap[c8][:, :, None] + np.arange(3)).reshape(len(cells), 24)
    KLf = grid.KL if KLf is None else KLf
    Ke = np.empty((len(cells), 576))
    did = grid.dim_id[cells]
    for d in np.unique(did):
        m = np.nonzero(did == d)[0]
        if coeffs is None:
            Ke[m] = lam[cells[m], None] * KLf[d].ravel() + mu[cells[m], None] * grid.KM[d].ravel()
        else:
            Ke[m] = coeffs[cells[m]] @ grid.KB[d].reshape(6, 576)
    K = _coo_to_csr(Ke, dofs, 3 * len(used))
    return K, nmap, used, cells, dofs


def face_node_areas(grid, cells, face, cn=None):
    """Four face nodes per cell and tributary area face/4 (fe/solver.py:165-174;
    rb/splitfe.py:94-102 with ``cn``)."""
    ax = "xyz".index(face[0])
    side = 1 if face[1] == "+" else 0
    sel = LOC[:, ax] == side
    conn = grid.cell_nodes if cn is None else cn
    nodes = conn[cells][:, sel]
    d = [grid.dx[grid.cI[cells]], grid.dy[grid.cJ[cells]], grid.dz[grid.cK[cells]]]
    o = [d[k] for k in range(3) if k != ax]
    area = np.repeat((o[0] * o[1] / 4.0)[:, None], 4, 1)
    return nodes.ravel(), area.ravel()


def rigid_modes(xyz, nmap, n, node_groups):
    """Columns = 3 translations + 3 rotations (r/20) per node group (fe/solver.py:177-199,
    rb/splitfe.py:105-120)."""
    rows, cols, vals = [], [], []
    col = 0
    for nodes in node_groups:
        nodes = np.asarray(nodes)
        ln = nmap[nodes]
        ok = ln >= 0
        nodes, ln = nodes[ok], ln[ok]
        if len(nodes) < 4:
            continue
        X = xyz[nodes]
        r = (X - X.mean(0)) / 20.0
        for d in range(3):
            rows.append(3 * ln + d); cols.append(np.full(len(ln), col)); vals.append(np.ones(len(ln)))
            col += 1
        rot = (((1, -r[:, 2]), (2, r[:, 1])), ((0, r[:, 2]), (2, -r[:, 0])), ((0, -r[:, 1]), (1, r[:, 0])))
        for pair in rot:
            for d, v in pair:
                rows.append(3 * ln + d); cols.append(np.full(len(ln), col)); vals.append(v)
            col += 1
    if col == 0:
        return sp.csr_matrix((n, 0))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, col))


# ---------------------------------------------------------------------------------------------
# fe/gmg.py
def p1d(coords):
    """1-D linear prolongation from nodes 0,2,4,... plus the last node (fe/gmg.py:7-26)."""
    coords = np.asarray(coords, float)
    n = len(coords)
    cidx = list(range(0, n, 2))
    if cidx[-1] != n - 1:
        cidx.append(n - 1)
    cidx = np.asarray(cidx)
    rows, cols, vals = [], [], []
    ci_of = {int(i): k for k, i in enumerate(cidx)}
    for i in range(n):
        if i in ci_of:
            rows.append(i); cols.append(ci_of[i]); vals.append(1.0)
            continue
        b = int(np.searchsorted(cidx, i)); a = b - 1
        w = (coords[i] - coords[cidx[a]]) / (coords[cidx[b]] - coords[cidx[a]])
        rows += [i, i]; cols += [a, b]; vals += [1.0 - w, w]
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, len(cidx))), coords[cidx]


def block_diag_inverse(A):
    """Inverse 3x3 diagonal blocks; blocks with |det| < 1e-30 become I (fe/gmg.py:29-44)."""
    A = A.tocoo()
    n = A.shape[0] // 3
    keep = (A.row // 3) == (A.col // 3)
    D = np.zeros((n, 3, 3))
    np.add.at(D, (A.row[keep] // 3, A.row[keep] % 3, A.col[keep] % 3), A.data[keep])
    bad = np.abs(np.linalg.det(D)) < 1e-30
    D[bad] = np.eye(3)
    return np.linalg.inv(D)


class Level:
    """One multigrid level: block-Jacobi M, spectral bound of M A by 12 power steps from a
    seed-0 normal vector, Chebyshev degree-3 smoothing (fe/gmg.py:47-76).  ``smoother='jacobi'``
    adds the damped block-Jacobi sweeps the benchmark configurations ask for."""

    def __init__(self, A, smoother="chebyshev", omega=0.6, sweeps=2, degree=3):
        self.A = A.tocsr()
        self.Dinv = block_diag_inverse(self.A)
        self.smoother, self.omega, self.sweeps, self.degree = smoother, omega, sweeps, degree
        if smoother == "chebyshev":
            x = np.random.default_rng(0).standard_normal(self.A.shape[0])
            lam = 1.0
            for _ in range(12):
                y = self.M(self.A @ x)
                lam = np.linalg.norm(y) / np.linalg.norm(x)
                x = y / np.linalg.norm(y)
            self.lmax = 1.1 * lam
            self.lmin = self.lmax / 30.0

    def M(self, r):
        return np.einsum("nij,nj->ni", self.Dinv, r.reshape(-1, 3)).ravel()

    def smooth(self, x, b):
        if self.smoother == "jacobi":
            for _ in range(self.sweeps):
                x = x + self.omega * self.M(b - self.A @ x)
            return x
        theta = 0.5 * (self.lmax + self.lmin)
        delta = 0.5 * (self.lmax - self.lmin)
        sigma = theta / delta
        rho = 1.0 / sigma
        r = b - self.A @ x
        d = self.M(r) / theta
        for k in range(self.degree):
            x = x + d
            if k == self.degree - 1:
                break
            r = r - self.A @ d
            rho_next = 1.0 / (2.0 * sigma - rho)
            d = rho_next * rho * d + (2.0 * rho_next / delta) * self.M(r)
            rho = rho_next
        return x


def prolongation(coords, act_dofs, axes=7):
    """Level prolongation P[act_dofs][:, keep] and next coords (fe/gmg.py:91-96).  ``axes``: bit
    mask of the coarsened axes (7 = all, the reference); an axis left out keeps its nodes
    (identity prolongation) -- the semi-coarsening option of the GPU build."""
    def p1(c, on):
        if on:
            return p1d(c)
        c = np.asarray(c, float)
        return sp.identity(len(c), format="csr"), c
    Px, cx = p1(coords[0], axes & 1); Py, cy = p1(coords[1], axes & 2); Pz, cz = p1(coords[2], axes & 4)
    Pn = sp.kron(Pz, sp.kron(Py, Px)).tocsr()
    Pd = sp.kron(Pn, sp.identity(3)).tocsr()[act_dofs]
    keep = np.nonzero(np.asarray(abs(Pd).sum(0)).ravel() > 0)[0]
    return Pd[:, keep].tocsr(), keep, (cx, cy, cz)


class GMG:
    """Galerkin geometric multigrid (fe/gmg.py:79-117): coarsen until <= min_dofs DOFs or
    max_levels levels; SuperLU coarse solve."""

    def __init__(self, A, used_nodes, xe, ye, ze, min_dofs=6000, max_levels=6, smoother="chebyshev",
                 omega=0.6, sweeps=2, levels=None, last_axes=7):
        self.levels, self.P, self.keeps = [], [], []
        act = (3 * np.asarray(used_nodes)[:, None] + np.arange(3)).ravel()
        coords = tuple(np.asarray(e, float) for e in (xe, ye, ze))
        Acur = A.tocsr()
        nlev = max_levels if levels is None else levels
        for lev in range(nlev):
            self.levels.append(Level(Acur, smoother, omega, sweeps))
            last = (lev == nlev - 1) if levels is not None else (Acur.shape[0] <= min_dofs)
            if last:
                break
            Pd, keep, coords_next = prolongation(coords, act)
            if levels is None and last_axes != 7 and len(keep) <= min_dofs:
                # the next level is the coarsest: coarsen only the axes in last_axes if that level
                # is still within min_dofs (no reference counterpart; the reference uses all axes)
                Ps, ks, cs = prolongation(coords, act, last_axes)
                if len(ks) <= min_dofs:
                    Pd, keep, coords_next = Ps, ks, cs
            coords = coords_next
            Acur = (Pd.T @ Acur @ Pd).tocsr()
            self.P.append(Pd)
            self.keeps.append(keep)
            act = keep
        self.coarse_A = Acur
        self.coarse = spla.splu(Acur.tocsc(), permc_spec="MMD_AT_PLUS_A")

    def vcycle(self, b, lev=0):
        if lev == len(self.levels) - 1:
            return self.coarse.solve(b)
        L = self.levels[lev]
        x = L.smooth(np.zeros_like(b), b)
        r = b - L.A @ x
        x = x + self.P[lev] @ self.vcycle(self.P[lev].T @ r, lev + 1)
        return L.smooth(x, b)


def pcg_gmg(A, b, used_nodes, grid_edges, tol=1e-7, maxiter=500, x0=None, A_op=None, Z=None, **mg_kw):
    """scipy CG preconditioned by the V-cycle, optional additive rigid-mode correction
    Z pinv(sym(Z^T A Z)) Z^T (fe/gmg.py:120-137)."""
    mg = GMG(A, used_nodes, *grid_edges, **mg_kw)
    n = A.shape[0]
    count = [0]

    def cb(_):
        count[0] += 1
    prec = mg.vcycle
    if Z is not None:
        Zd = Z.toarray() if sp.issparse(Z) else np.asarray(Z)
        Aeff = A if A_op is None else A_op
        E = Zd.T @ np.column_stack([Aeff @ Zd[:, i] for i in range(Zd.shape[1])])
        Einv = np.linalg.pinv(0.5 * (E + E.T))
        prec = lambda r: mg.vcycle(r) + Zd @ (Einv @ (Zd.T @ r))
    M = spla.LinearOperator((n, n), matvec=prec)
    x, flag = spla.cg(A if A_op is None else A_op, b, x0=x0, rtol=tol, maxiter=maxiter, M=M, callback=cb)
    return x, dict(iters=count[0], flag=flag, levels=len(mg.levels), mg=mg)


def pcg_cadence(A, b, prec, tol, maxiter, x0=None):
    """The hand-rolled PCG of rb/gpusolve.py:229-254: residual norm checked when it < 5 or
    it % 5 == 0; threshold tol*||b|| (||b||=0 -> 1).  Returns x, iters, flag, relres."""
    x = np.zeros_like(b) if x0 is None else np.asarray(x0, float).copy()
    r = b - A @ x
    bn = np.linalg.norm(b)
    thr = tol * (bn if bn > 0 else 1.0)
    z = prec(r); p = z.copy(); rz = r @ z
    it, flag = 0, 1
    rn = np.linalg.norm(r)
    if rn <= thr:
        flag = 0
    while flag and it < maxiter:
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        it += 1
        if it % 5 == 0 or it < 5:
            rn = np.linalg.norm(r)
            if rn <= thr:
                flag = 0
                break
        z = prec(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    if flag:
        rn = np.linalg.norm(r)
        flag = 0 if rn <= thr else maxiter
    return x, it, flag, rn / (bn if bn > 0 else 1.0)


# ---------------------------------------------------------------------------------------------
# fe/solver.py:202-254
def solve(grid, lam, mu, active, springs, loads, method="cg", tol=1e-8, x0=None, maxiter=20000, penalties=(),
          coeffs=None, node_groups=None, U0=None, **mg_kw):
    K, nmap, used, cells, dofs = assemble(grid, lam, mu, active, coeffs)
    n = K.shape[0]
    if U0 is not None:
        x0 = np.nan_to_num(U0[used]).ravel()
    f = np.zeros(n); kd = np.zeros(n)
    for nodes, area, kvec in springs:
        ln = nmap[nodes]; ok = ln >= 0
        for c in range(3):
            np.add.at(kd, 3 * ln[ok] + c, area[ok] * kvec[c])
    for nodes, area, t in loads:
        ln = nmap[nodes]; ok = ln >= 0
        for c in range(3):
            np.add.at(f, 3 * ln[ok] + c, area[ok] * t[c])
    K = K + sp.diags(kd)
    for nodes, weights, comp, k in penalties:                      # fe/solver.py:224-232
        ln = nmap[nodes]; ok = ln >= 0
        w = np.zeros(n)
        np.add.at(w, 3 * ln[ok] + comp, weights[ok])
        w /= w.sum()
        nz = np.nonzero(w)[0]
        K = K + sp.csr_matrix((k * np.outer(w[nz], w[nz]).ravel(), (np.repeat(nz, len(nz)), np.tile(nz, len(nz)))),
                              shape=K.shape)
    info = {}
    if method == "none":                                            # assembled system only
        info.update(ndof=n, K=K, f=f, used=used, nmap=nmap)
        return None, info
    if method == "gmg":
        Z = rigid_modes(grid.node_xyz, nmap, n, node_groups) if node_groups else None
        u, inf = pcg_gmg(K, f, used, (grid.xe, grid.ye, grid.ze), tol=tol, x0=x0, Z=Z, maxiter=maxiter, **mg_kw)
        info.update(inf)
    elif method == "direct":
        u = spla.spsolve(K.tocsc(), f, permc_spec="MMD_AT_PLUS_A")
        info["iters"] = 0
    else:
        M = sp.diags(1.0 / K.diagonal())
        count = [0]

        def cb(_):
            count[0] += 1
        u, flag = spla.cg(K, f, x0=x0, rtol=tol, maxiter=maxiter, M=M, callback=cb)
        info["iters"] = count[0]; info["flag"] = flag
    U = np.full((grid.nnode, 3), np.nan)
    U[used] = u.reshape(-1, 3)
    info.update(ndof=n, K=K, f=f, used=used, nmap=nmap)
    return U, info


def cell_stress(grid, lam, mu, U, cells, coeffs=None, cn=None):
    """Centre strain, stress and von Mises (fe/solver.py:257-272; rb/splitfe.py:161-175 with
    ``cn``, which also maps NaN to 0)."""
    conn = grid.cell_nodes if cn is None else cn
    ue = U[conn[cells]].reshape(len(cells), 24)
    if cn is not None:
        ue = np.nan_to_num(ue)
    eps = np.einsum("nij,nj->ni", grid.Bc[grid.dim_id[cells]], ue)
    if coeffs is None:
        tr = eps[:, :3].sum(1)
        sig = np.empty_like(eps)
        sig[:, :3] = lam[cells, None] * tr[:, None] + 2 * mu[cells, None] * eps[:, :3]
        sig[:, 3:] = mu[cells, None] * eps[:, 3:]
    else:
        D = np.einsum("nk,kij->nij", coeffs[cells], ortho_patterns())
        sig = np.einsum("nij,nj->ni", D, eps)
    return sig, von_mises(sig), eps


def von_mises(sig):
    sx, sy, sz, txy, tyz, txz = sig.T
    return np.sqrt(0.5 * ((sx - sy) ** 2 + (sy - sz) ** 2 + (sz - sx) ** 2) + 3 * (txy ** 2 + tyz ** 2 + txz ** 2))


# ---------------------------------------------------------------------------------------------
# rb/splitfe.py
def split_connectivity(grid, active, part, bone_share):
    """Loop form (rb/splitfe.py:17-55): a node touched by cells of >= 2 bones gets one copy per
    extra bone (ids nnode + rank in sorted (node, bone) order); part-0 cells attach to the bone
    with the largest share in the cell, else to the bone with most cells at the node."""
    cells = np.nonzero(active)[0]
    c8 = grid.cell_nodes[cells]
    pc = part[cells]
    nb = bone_share.shape[1]
    touch = np.zeros((grid.nnode, nb), np.int64)
    for l in range(8):
        np.add.at(touch, (c8[:, l], pc), 1)
    touch[:, 0] = 0
    split = np.nonzero((touch > 0).sum(1) >= 2)[0]
    new_id, extra = {}, []
    nxt = grid.nnode
    first = {}
    for n in split:
        bones = np.nonzero(touch[n])[0]
        first[int(n)] = int(bones[0])
        for b in bones[1:]:
            new_id[(int(n), int(b))] = nxt
            extra.append(int(n))
            nxt += 1
    orig = np.concatenate([np.arange(grid.nnode), np.asarray(extra, np.int64)])
    cn = grid.cell_nodes.copy()
    is_split = np.zeros(grid.nnode, bool); is_split[split] = True
    for r in range(len(cells)):
        for l in range(8):
            n = int(c8[r, l])
            if not is_split[n]:
                continue
            c = cells[r]
            b = int(pc[r])
            if b == 0:
                cand = np.nonzero(touch[n])[0]
                sh = bone_share[c, cand]
                b = int(cand[np.argmax(sh)]) if sh.max() > 0 else int(cand[np.argmax(touch[n, cand])])
            if b != first[n]:
                cn[c, l] = new_id[(n, b)]
    return cn, orig, len(split)


def split_connectivity_fast(grid, active, part, bone_share):
    """Vectorised form (rb/splitfe.py:261-305), same result as the loop form."""
    cells = np.nonzero(active)[0]
    c8 = grid.cell_nodes[cells]
    pc = part[cells]
    nb = bone_share.shape[1]
    bone_cells = pc > 0
    key = c8[bone_cells].astype(np.int64).ravel() * nb + np.repeat(pc[bone_cells], 8)
    ukey, cnt = np.unique(key, return_counts=True)
    unode, ubone = ukey // nb, ukey % nb
    first = np.ones(len(ukey), bool)
    first[1:] = unode[1:] != unode[:-1]
    npair = np.bincount(unode, minlength=grid.nnode)
    is_split = npair >= 2
    extra = ~first & is_split[unode]
    newid = np.full(len(ukey), -1, np.int64)
    newid[extra] = grid.nnode + np.arange(int(extra.sum()))
    orig = np.concatenate([np.arange(grid.nnode), unode[extra]])
    pos0 = np.searchsorted(unode, np.arange(grid.nnode))
    cn = grid.cell_nodes.copy()
    rows, ls = np.nonzero(is_split[c8])
    if len(rows):
        n = c8[rows, ls]
        c = cells[rows]
        b = pc[rows].astype(np.int64)
        z = np.nonzero(b == 0)[0]
        if len(z):
            nz_ = n[z]
            k = npair[nz_]
            best = np.full(len(z), -1, np.int64); bsh = np.full(len(z), -1.0)
            bcnt = np.full(len(z), -1); bcnt_b = np.full(len(z), -1, np.int64)
            for j in range(int(k.max())):
                ok = j < k
                idx = np.nonzero(ok)[0]
                pidx = pos0[nz_[ok]] + j
                bb = ubone[pidx]
                sh = bone_share[c[z][ok], bb]
                ct = cnt[pidx]
                up = sh > bsh[ok]
                best[idx[up]] = bb[up]; bsh[idx[up]] = sh[up]
                up2 = ct > bcnt[ok]
                bcnt_b[idx[up2]] = bb[up2]; bcnt[idx[up2]] = ct[up2]
            b[z] = np.where(bsh > 0, best, bcnt_b)
        pk = np.searchsorted(ukey, n.astype(np.int64) * nb + b)
        nid = newid[pk]
        sel = nid >= 0
        cn[c[sel], ls[sel]] = nid[sel]
    return cn, orig, int(is_split.sum())


def assemble_cn(grid, lam, mu, cells_or_active, cn, sri=True):
    """rb/splitfe.py:72-91 (SRI volumetric term when ``sri``)."""
    active = np.zeros(grid.ncell, bool)
    cells = np.asarray(cells_or_active)
    if cells.dtype == bool:
        active = cells
    else:
        active[cells] = True
    K, nmap, used, _, _ = assemble(grid, lam, mu, active, KLf=grid.KL_sri if sri else grid.KL, cn=cn)
    return K, nmap, used


def solve_cn(grid, lam, mu, active, cn, orig, fixed, loads, extra=None, node_groups=None, tol=1e-6, maxiter=800,
             method="gmg", sri=True, **mg_kw):
    """rb/splitfe.py:123-158: ground springs k = 1e6 * area on 3 components, loads, extra(nmap, n)
    stiffness, optional rigid modes, then pcg_gmg with the copies' original grid nodes."""
    K, nmap, used, cells, _ = assemble(grid, lam, mu, active, KLf=grid.KL_sri if sri else grid.KL, cn=cn)
    n = K.shape[0]
    f = np.zeros(n); kd = np.zeros(n)
    nb_, ab_ = fixed
    ln = nmap[nb_]
    for c in range(3):
        np.add.at(kd, 3 * ln + c, ab_ * 1e6)
    for nodes, area, t in loads:
        ln = nmap[nodes]
        for c in range(3):
            np.add.at(f, 3 * ln + c, area * t[c])
    K = K + sp.diags(kd)
    if extra is not None:
        K = K + extra(nmap, n)
    xyz = grid.node_xyz[orig]
    Z = rigid_modes(xyz, nmap, n, node_groups) if node_groups else None
    info = {}
    if method == "none":                                            # assembled system only
        info.update(ndof=n, K=K, f=f, used=used, nmap=nmap, Z=Z)
        return None, info, None
    if method == "direct":
        u = spla.spsolve(K.tocsc(), f, permc_spec="MMD_AT_PLUS_A")
        info.update(iters=0, flag=0)
    else:
        u, info = pcg_gmg(K, f, orig[used], (grid.xe, grid.ye, grid.ze), tol=tol, Z=Z, maxiter=maxiter, **mg_kw)
    U = np.full((len(orig), 3), np.nan)
    U[used] = u.reshape(-1, 3)
    R = np.zeros((len(orig), 3))
    np.add.at(R, nb_, -(ab_[:, None] * 1e6) * U[nb_])
    info.update(ndof=n, K=K, f=f, used=used, nmap=nmap)
    return U, info, R


def element_forces_cn(grid, lam, mu, U, cells, cn, sri=True):
    """Per-cell nodal forces K_e u_e (rb/splitfe.py:178-193)."""
    ue = np.nan_to_num(U[cn[cells]].reshape(len(cells), 24))
    KLf = grid.KL_sri if sri else grid.KL
    out = np.zeros((len(cells), 24))
    did = grid.dim_id[cells]
    for d in np.unique(did):
        m = did == d
        out[m] = lam[cells[m], None] * (ue[m] @ KLf[d].T) + mu[cells[m], None] * (ue[m] @ grid.KM[d].T)
    return out.reshape(len(cells), 8, 3)


# ---------------------------------------------------------------------------------------------
# rb/replica3.py:68-86
def spring_patches(xyz, springs, bone_nodes, used):
    """rb/replica3.py:74-80: the patch of each spring end: the end's bone's used nodes within
    max(4 mm, distance of the 8th nearest).  Returns [(patch of a, patch of b)]."""
    out = []
    for s in springs:
        pp = []
        for p, key in ((s["a"], "A"), (s["b"], "B")):
            cand = np.asarray(bone_nodes[s[key]])
            cand = cand[used[cand]]
            dist = np.linalg.norm(xyz[cand] - np.asarray(p, float), axis=1)
            pp.append(cand[dist <= max(4.0, np.sort(dist)[min(7, len(dist) - 1)])])
        out.append(tuple(pp))
    return out


def spring_matrix_cn(xyz, springs, bone_nodes):
    """Ligament springs: K += k w w^T with w = +d/n_A on the end-A patch and -d/n_B on the end-B
    patch (patch = bone nodes within max(4 mm, 8th-nearest distance))."""
    def build(nmap, n):
        rows, cols, vals = [], [], []
        patches = spring_patches(xyz, springs, bone_nodes, nmap >= 0)
        for s, (pa, pb) in zip(springs, patches):
            a, b = np.asarray(s["a"], float), np.asarray(s["b"], float)
            d = (b - a) / np.linalg.norm(b - a)
            w = np.zeros(n)
            for sel, sign in ((pa, 1.0), (pb, -1.0)):
                ln = nmap[sel]
                for c in range(3):
                    np.add.at(w, 3 * ln + c, sign * d[c] / len(ln))
            nz = np.nonzero(w)[0]
            rows.append(np.repeat(nz, len(nz))); cols.append(np.tile(nz, len(nz)))
            vals.append(s["k"] * np.outer(w[nz], w[nz]).ravel())
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return build


# ---------------------------------------------------------------------------------------------
# rb/rbsim.py:110-140
def sample_fe(g, cn, active, VM, SIG, Urel, grid_ml, sub=2):
    """Average VM, stress and displacement of each analysis cell from sub^3 points in the FE
    cells; inactive FE cells are excluded.  Returns vm_c, u_c, cover, sig_c."""
    oo = (np.arange(sub) + 0.5) / sub
    X = (grid_ml.xe[:-1, None] + grid_ml.dx[:, None] * oo).ravel()
    Y = (grid_ml.ye[:-1, None] + grid_ml.dy[:, None] * oo).ravel()
    Z = (grid_ml.ze[:-1, None] + grid_ml.dz[:, None] * oo).ravel()

    def locate(e, d, v, n):
        idx = np.clip(np.searchsorted(e, v) - 1, 0, n - 1)
        t = np.clip((v - e[idx]) / d[idx], 0, 1)
        inside = (v >= e[0]) & (v <= e[-1])
        return idx, t, inside
    i, tx, inx = locate(g.xe, g.dx, X, g.nx)
    j, ty, iny = locate(g.ye, g.dy, Y, g.ny)
    k, tz, inz = locate(g.ze, g.dz, Z, g.nz)
    I, J, K = np.meshgrid(i, j, k, indexing="ij")
    c = ((I * g.ny + J) * g.nz + K).ravel()
    ok = (inx[:, None, None] & iny[None, :, None] & inz[None, None, :]).ravel() & active[c]
    TX, TY, TZ = (a.ravel() for a in np.meshgrid(tx, ty, tz, indexing="ij"))
    u = np.zeros((len(c), 3))
    for l in range(8):
        a, b, cc = LOC[l]
        w = (TX if a else 1 - TX) * (TY if b else 1 - TY) * (TZ if cc else 1 - TZ)
        u += w[:, None] * np.nan_to_num(Urel[cn[c, l]])
    vm = np.where(ok, VM[c], 0.0)
    sg = np.where(ok[:, None], SIG[c], 0.0)
    nx, ny, nz = grid_ml.shape

    def blocksum(a):
        a = a.reshape(nx, sub, ny, sub, nz, sub, *a.shape[1:])
        a = a.transpose(0, 2, 4, 1, 3, 5, *range(6, a.ndim))
        return a.reshape(nx * ny * nz, sub ** 3, *a.shape[6:]).sum(1)
    okc = blocksum(ok.astype(np.float64))
    den = np.maximum(okc, 1)
    vm_c = blocksum(vm) / den
    u_c = blocksum(u * ok[:, None]) / den[:, None]
    sig_c = blocksum(sg) / den[:, None]
    return vm_c.astype(np.float32), u_c.astype(np.float32), okc / sub ** 3, sig_c.astype(np.float32)


# rb/rbsim.py:143-161
def sample_nodes(g, cn, active, Urel, grid_ml):
    """Relative displacement at every analysis-grid node (grid_ml.node_xyz order): trilinear in the
    FE cell holding the node (clip(searchsorted - 1)), or, when that cell is inactive, in the
    nearest active cell of the Euclidean feature transform of the inactive mask in cell-index units
    (scipy.ndimage.distance_transform_edt, rb/rbsim.py:152), local coordinates clamped to [0, 1].
    Returns float32 (nnode_ml, 3)."""
    from scipy import ndimage
    P = grid_ml.node_xyz
    held = [np.clip(np.searchsorted(e, P[:, a]) - 1, 0, n - 1)
            for a, (e, n) in enumerate(((g.xe, g.nx), (g.ye, g.ny), (g.ze, g.nz)))]
    _, feat = ndimage.distance_transform_edt(~np.asarray(active, bool).reshape(g.shape), return_indices=True)
    near = [feat[a][held[0], held[1], held[2]] for a in range(3)]
    c = (near[0] * g.ny + near[1]) * g.nz + near[2]
    t = [np.clip((P[:, a] - e[near[a]]) / d[near[a]], 0, 1)
         for a, (e, d) in enumerate(((g.xe, g.dx), (g.ye, g.dy), (g.ze, g.dz)))]
    u = np.zeros((len(P), 3))
    for l in range(8):
        w = np.ones(len(P))
        for a in range(3):
            w = w * (t[a] if LOC[l, a] else 1 - t[a])
        u += w[:, None] * np.nan_to_num(Urel[cn[c, l]])
    return u.astype(np.float32)


def nearest_active_bruteforce(active, shape, cells):
    """Nearest active cell of every listed cell by enumeration: all active cells at the smallest
    squared index distance, and the one of them with the smallest (k, j, i)."""
    A = np.argwhere(np.asarray(active, bool).reshape(shape))
    Q = np.stack(np.unravel_index(np.asarray(cells), shape), 1)
    d2 = ((Q[:, None, :] - A[None, :, :]) ** 2).sum(-1)
    best = d2.min(1)
    order = np.lexsort((A[:, 0], A[:, 1], A[:, 2]))
    first = np.argmax(d2[:, order] == best[:, None], axis=1)
    return A[order][first], best


def rel_l2(a, b):
    """Relative L2 error ||a-b|| / ||b|| (ml/evaluate.py:93,96)."""
    a = np.asarray(a, float).ravel(); b = np.asarray(b, float).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------------------------------------
# Lattice materials (the engine's voxel preparation).  A lattice V holds, per cell in reference
# order, ns^3 sub-sample body labels L (ncell, ns^3) and a body -> class map cls.
# rb/rbgeom.py:45-49: tissue classes and their (E, nu); cartilage and soft tissue are overridden
# by materials()'s arguments.
LATTICE_CLASSES = ("void", "cortical", "cancellous", "cartilage", "fascia", "soft", "plate", "fibrous", "fat", "gap")
LATTICE_MAT = {"cortical": (7300.0, 0.3), "cancellous": (100.0, 0.3), "cartilage": (10.0, 0.4),
               "fascia": (350.0, 0.4), "soft": (1.15, 0.49), "plate": (17000.0, 0.1), "fibrous": (60.0, 0.4),
               "fat": (1.15, 0.45), "gap": (0.01, 0.3)}


def bone_groups(cls, roles):
    """rb/replica2.py:34-46: group ids 1.. for the named bones (in roles order), then every other
    bone body (class cortical or cancellous, not bonded_extra) on its own, in label order."""
    grp, names = {}, {}
    gid = 1
    for name, bodies in roles["bones"].items():
        for b in bodies:
            grp[b] = gid
        names[gid] = name
        gid += 1
    bonded = set(roles.get("bonded_extra", ()))
    for b in sorted(cls):
        if cls[b] in (1, 2) and b not in grp and b not in bonded:
            grp[b] = gid
            names[gid] = f"body{b}"
            gid += 1
    return grp, names


def bone_share(L, cls, roles):
    """rb/replica3.py:17-29: fraction of each cell's sub-samples in every bone group (float32,
    column 0 unused)."""
    grp, names = bone_groups(cls, roles)
    nlab = max(max(cls), int(L.max())) + 1
    lutg = np.zeros(nlab, np.int64)
    for b, g in grp.items():
        lutg[b] = g
    Gb = lutg[L]
    S = np.zeros((L.shape[0], max(names) + 1), np.float32)
    for g in range(1, max(names) + 1):
        S[:, g] = (Gb == g).mean(1)
    return S, names


def lattice_materials(L, cls, roles, dims, cart_E=10.0, nu_soft=0.49):
    """rb/rbfe.py:75-134 (mixing rule iface3 of replica2 with fibrous layers, fat and gaps).
    dims = (nx, ny, nz); rows of L in reference cell order.  Returns E, nu, active, F, ser."""
    from scipy import ndimage
    nlab = max(max(cls), int(L.max())) + 1
    lut = np.zeros(nlab, np.int64)
    for b, c in cls.items():
        lut[b] = c
    C = lut[L]
    F = np.stack([(C == k).mean(1) for k in range(len(LATTICE_CLASSES))], 1)
    mat = dict(LATTICE_MAT)
    mat["cartilage"] = (cart_E, 0.4)
    mat["soft"] = (1.15, nu_soft)
    E_k = np.array([1.0] + [mat[c][0] for c in LATTICE_CLASSES[1:]])
    nu_k = np.array([0.3] + [mat[c][1] for c in LATTICE_CLASSES[1:]])
    lam_k, mu_k = lame(E_k, nu_k)
    M_k = lam_k + 2 * mu_k
    H, LAY, SOFT, GAPC = [1, 2, 4], [3, 7], [5, 8], [9]
    PH = H + LAY + SOFT
    occ = F[:, PH].sum(1)
    hard, lay, fgap = F[:, H].sum(1), F[:, LAY].sum(1), F[:, GAPC].sum(1)
    with np.errstate(divide="ignore", invalid="ignore"):
        # Voigt average over the occupied phases
        lamV = (F[:, PH] * lam_k[PH]).sum(1) / occ
        muV = (F[:, PH] * mu_k[PH]).sum(1) / occ
        # hard phases in series with joint layers / soft tissue (Reuss on constrained and shear moduli)
        lam_h = (F[:, H] * lam_k[H]).sum(1) / hard
        mu_h = (F[:, H] * mu_k[H]).sum(1) / hard
        hM = np.where(hard > 0, hard / (lam_h + 2 * mu_h), 0.0)
        hG = np.where(hard > 0, hard / mu_h, 0.0)
        gapM = (F[:, GAPC] / M_k[GAPC]).sum(1)
        gapG = (F[:, GAPC] / mu_k[GAPC]).sum(1)
        layM = (F[:, LAY] / M_k[LAY]).sum(1) + gapM
        layG = (F[:, LAY] / mu_k[LAY]).sum(1) + gapG
        fch = lay + hard + fgap
        occ_g = occ + fgap
        M_ch, G_ch = fch / (layM + hM), fch / (layG + hG)
        lam_soft, mu_soft = (F[:, SOFT] * lam_k[SOFT]).sum(1), (F[:, SOFT] * mu_k[SOFT]).sum(1)
        lam_c = (fch * (M_ch - 2 * G_ch) + lam_soft) / occ_g
        mu_c = (fch * G_ch + mu_soft) / occ_g
        softM, softG = (F[:, SOFT] / M_k[SOFT]).sum(1), (F[:, SOFT] / mu_k[SOFT]).sum(1)
        M_s = occ_g / (softM + hM + gapM)
        G_s = occ_g / (softG + hG + gapG)
        lam_s, mu_s = M_s - 2 * G_s, G_s
    # joint cells: within one cell (3x3x3) of two different bone groups
    grp, _ = bone_groups(cls, roles)
    lutg = np.zeros(nlab, np.int64)
    for b, g in grp.items():
        lutg[b] = g
    Gb = lutg[L]
    count = np.zeros(L.shape[0], np.int64)
    for g in np.unique(Gb[Gb > 0]):
        pres = (Gb == g).any(1).reshape(dims)
        count += ndimage.binary_dilation(pres, np.ones((3, 3, 3), bool)).ravel()
    cart = lay > 0
    ser = cart | (count >= 2)
    lamX = np.where(cart, lam_c, np.where(ser, lam_s, lamV))
    muX = np.where(cart, mu_c, np.where(ser, mu_s, muV))
    with np.errstate(divide="ignore", invalid="ignore"):
        E = muX * (3 * lamX + 2 * muX) / (lamX + muX)
        nu = lamX / (2 * (lamX + muX))
    active = np.where(ser, occ > 0.15, occ >= 0.5)
    E = np.where(active, E, 1.0)
    nu = np.where(active, np.clip(nu, 0.0, 0.4995), 0.3)
    return np.nan_to_num(E, nan=1.0), np.nan_to_num(nu, nan=0.3), active, F, ser


def lattice_part(S, F):
    """rb/rbfe.py:140-141: bone group of a cell holding >= 50 % of one bone and no joint layer."""
    cart = (F[:, 3] + F[:, 7]) > 0
    return np.where((S.max(1) >= 0.5) & ~cart, S.argmax(1), 0)


# ---------------------------------------------------------------------------------------------
# rb/rbfe.py:146-178  (ground contact, leg loads, bone node sets, spring ends, rigid groups)
def bone_node_sets(cn, part, active, keys):
    """rb/rbfe.py:166-169,178: for every part id k in ``keys``: the sorted unique nodes of the active
    cells of part k.  Returns {k: nodes}."""
    cells = np.arange(len(part))
    return {k: np.unique(cn[cells[(part == k) & active]].ravel()) for k in keys}


def part_cell_counts(part, active, keys):
    """Active cells per part id (the rigid-group threshold of rb/rbfe.py:178 is > 3)."""
    return {k: int(((part == k) & active).sum()) for k in keys}


def ground_contact(g, cn, active, contact):
    """rb/rbfe.py:146-151: bottom-layer active cells with contact fraction cf > 0 (cf per (i, j)
    column, ``contact`` shaped (nx, ny)); their z- face nodes, areas face/4 times cf."""
    cells = np.arange(g.ncell)
    cf = np.asarray(contact).ravel()[g.cI * g.ny + g.cJ]
    fc = cells[(g.cK == 0) & active & (cf > 0)]
    nodes, area = face_node_areas(g, fc, "z-", cn)
    return nodes, area * np.repeat(cf[fc], 4)


def leg_loads(g, cn, orig, active, top_fraction, forces, load, load_line):
    """rb/rbfe.py:152-163: for every loaded bone, the z+ faces of the top-layer active cells with
    more than half of the bone's sub-samples (``top_fraction[name]``, per cell) carry a linear
    pressure tz = -Fz/sum(w) (1 + dX ab) whose resultant is -Fz along z through the load line:
    ab = I^-1 sum(w) (line - centroid), I the area second moment of the loaded nodes."""
    cells = np.arange(g.ncell)
    top = cells[(g.cK == g.nz - 1) & active]
    out = []
    for name, Fz in forces.items():
        Fz = Fz * load
        c = top[top_fraction[name][top] > 0.5]
        n_, a_ = face_node_areas(g, c, "z+", cn)
        X = g.node_xyz[orig[n_]][:, :2]
        w = a_
        xb = (w[:, None] * X).sum(0) / w.sum()
        dX = X - xb
        I = np.einsum("n,ni,nj->ij", w, dX, dX)
        e = np.array(load_line[name]) - xb
        ab = np.linalg.solve(I, w.sum() * e)
        tz = -Fz / w.sum() * (1 + dX @ ab)
        out.append((n_, a_ * tz, np.array([0.0, 0.0, 1.0])))
    return out


def spring_end_bones(springs, bone_nodes, xyz):
    """rb/rbfe.py:173-175: each spring end goes to the bone whose nearest node is nearest (the
    first such bone in ``bone_nodes`` order; bones without nodes are skipped).  Brute force: the
    reference queries one KD-tree per bone, whose nearest-neighbour distance is the same minimum."""
    names = [nm for nm, nd in bone_nodes.items() if len(nd)]
    out = []
    for s in springs:
        ends = []
        for p in (s["a"], s["b"]):
            p = np.asarray(p, float)
            best, bd = None, None
            for nm in names:
                d = np.sqrt(((xyz[bone_nodes[nm]] - p) ** 2).sum(1)).min()
                if bd is None or d < bd:
                    best, bd = nm, d
            ends.append(best)
        out.append(tuple(ends))
    return out


# ---------------------------------------------------------------------------------------------
# rb/rbfe.py:188-219 (region labels, forces on the talus / calcaneus, ground reactions by zone) and
# rb/rbsim.py:99-107,164-203 (rigid fit, effect descriptors)
ENGINE_CART = {"ankle": 5, "syndesmosis": 6, "talofibular": 7, "mid": 8, "tn": 9, "ant": 10, "cc": 11, "post": 12}
COAL, FIB, INTERP, FUSE = 101, 102, 103, 104


def body_fraction(L, bodies):
    """rb/rbfe.py:138 bf: fraction of a cell's sub-samples whose body is in ``bodies``."""
    return np.isin(L, list(bodies)).mean(1)


def region_labels(F, L, cart=ENGINE_CART):
    """rb/rbfe.py:190-201: region of every cell, later rules overriding earlier ones: 'other', then
    the dominant class (soft, fascia, bone = cortical or cancellous), then cells holding any
    coalition/fusion, fibrous, interposition and finally any cartilage body.  Returns (lab, names)."""
    names = list(cart) + ["coalition", "fibrous", "fat", "soft", "fascia", "bone", "other"]
    lab = np.full(len(F), len(names) - 1, np.int16)
    dom = np.argmax(F[:, 1:], 1) + 1
    lab[dom == LATTICE_CLASSES.index("soft")] = names.index("soft")
    lab[dom == LATTICE_CLASSES.index("fascia")] = names.index("fascia")
    lab[np.isin(dom, [1, 2])] = names.index("bone")
    lab[body_fraction(L, (COAL, FUSE)) > 0] = names.index("coalition")
    lab[body_fraction(L, (FIB,)) > 0] = names.index("fibrous")
    lab[body_fraction(L, (INTERP,)) > 0] = names.index("fat")
    for key, b in cart.items():
        lab[body_fraction(L, (b,)) > 0] = names.index(key)
    return lab, names


def bone_forces(grid, lam, mu, U, cn, part, active, k, lab, names, nnode_total, sri=True):
    """rb/rbfe.py:202-210 without the final rounding: the force on bone part k from every region,
    minus the element forces (K_e u_e) that the active cells outside the bone put on the bone's
    nodes, summed by the region of the cell.  Returns {name: [Fx, Fy, Fz]} for the regions present."""
    cells = np.arange(grid.ncell)
    tgt = (part == k) & active
    tn = np.zeros(nnode_total, bool)
    tn[cn[cells[tgt]].ravel()] = True
    oc = cells[active & ~tgt]
    oc = oc[tn[cn[oc]].any(1)]
    fe = element_forces_cn(grid, lam, mu, U, oc, cn, sri=sri)
    fsum = -(fe * tn[cn[oc]][:, :, None]).sum(1)
    return {nm: list(fsum[lab[oc] == i].sum(0)) for i, nm in enumerate(names) if (lab[oc] == i).any()}


def ground_zones(R, x, scale, c0x):
    """rb/rbfe.py:211-215 without the final rounding: vertical reactions summed over the heel
    (x < xh), mid and fore (x >= xf) zones, xh = c0x + s (-40 - c0x), xf = c0x + s (20 - c0x)."""
    xh = c0x + scale * (-40.0 - c0x)
    xf = c0x + scale * (20.0 - c0x)
    return {"heel": float(R[x < xh, 2].sum()), "mid": float(R[(x >= xh) & (x < xf), 2].sum()),
            "fore": float(R[x >= xf, 2].sum())}


def rigid_fit(X, U):
    """rb/rbsim.py:99-107: least-squares translation t and small rotation w about the centroid c of
    X fitting U ~ t + w x (X - c).  Returns (t, w, c)."""
    c = X.mean(0)
    r = X - c
    A = np.zeros((3 * len(X), 6))
    A[0::3, 0] = 1; A[1::3, 1] = 1; A[2::3, 2] = 1
    A[0::3, 4] = r[:, 2]; A[0::3, 5] = -r[:, 1]
    A[1::3, 3] = -r[:, 2]; A[1::3, 5] = r[:, 0]
    A[2::3, 3] = r[:, 1]; A[2::3, 4] = -r[:, 0]
    sol, *_ = np.linalg.lstsq(A, U.ravel(), rcond=None)
    return sol[:3], sol[3:], c


def effects(VM, L, active, part, bnames, vol, U, xyz, bone_nodes, F_on_talus, ground_Fz, ground_total_Fz,
            cart=ENGINE_CART, dt=None, dc=None):
    """rb/rbsim.py:164-203: cartilage p99 / volume-weighted mean von Mises, talus load shares,
    calcaneus-relative talus translation and eversion, navicular drop, heel share, coalition p99 and
    max; tc_p99 (the talocalcaneal band) only when the template distances dt, dc of the cell centres
    to the talus and calcaneus are given (rb/rbsim.py:193-199).  xyz: positions of all nodes
    (split copies at their original node).  Returns (d, Urel)."""
    bf = lambda bodies: body_fraction(L, bodies)
    d = {}
    for key in ("ant", "mid", "post", "tn", "cc"):
        fb = bf((cart[key],))
        m = (fb >= 0.25) & active
        mm = (fb > 0) & active
        d[key + "_p99"] = float(np.percentile(VM[m], 99)) if m.sum() >= 3 else 0.0
        w = fb[mm] * vol[mm]
        d[key + "_mean"] = float((VM[mm] * w).sum() / w.sum()) if mm.any() else 0.0
    fz = {k: max(F_on_talus.get(k, [0, 0, 0])[2], 0.0) for k in ("ant", "mid", "post")}
    tot = sum(fz.values()) + 1e-9
    for k in fz:
        d["share_" + k] = float(fz[k] / tot)
    nodes_c = bone_nodes["calcaneus"]
    t, w, c = rigid_fit(xyz[nodes_c], np.nan_to_num(U[nodes_c]))
    Urel = np.nan_to_num(U) - (t[None, :] + np.cross(w[None, :], xyz - c[None, :]))
    nodes_t = bone_nodes["talus"]
    tt, wt, ct = rigid_fit(xyz[nodes_t], Urel[nodes_t])
    d["talus_dz"] = float(tt[2]); d["talus_dy"] = float(tt[1]); d["talus_evert"] = float(np.rad2deg(wt[0]))
    d["nav_drop"] = float(-np.mean(Urel[bone_nodes["navicular"], 2]))
    d["heel_share"] = float(ground_Fz["heel"] / max(ground_total_Fz, 1e-9))
    if dt is not None:
        kt = [i for i, n in bnames.items() if n == "talus"][0]
        kc = [i for i, n in bnames.items() if n == "calcaneus"][0]
        sub_ = bf((cart["post"], cart["mid"], cart["ant"], COAL, FUSE, FIB))
        band = active & (dt <= 6) & (dc <= 6) & ((part == kt) | (part == kc) | (sub_ > 0))
        d["tc_p99"] = float(np.percentile(VM[band], 99)) if band.sum() > 3 else 0.0
    cm = (bf((COAL,)) >= 0.25) & active
    d["coal_p99"] = float(np.percentile(VM[cm], 99)) if cm.sum() > 3 else 0.0
    d["coal_max"] = float(VM[cm].max()) if cm.sum() > 0 else 0.0
    return d, Urel
