"""Rectilinear voxel grid and Q1 hex element matrices (host side).

Same numbering and attributes as ``fe/solver.py:93-128`` (``Grid``): nodes x-fastest
``i + (nx+1)(j + (ny+1)k)``, cells z-fastest ``(i*ny + j)*nz + k``, local corner order
``l = a + 2b + 4c``.  Heavy per-cell/per-node arrays (``cell_nodes``, ``node_xyz`` ...) are built on
first access so that a 25M-cell grid costs nothing until a caller asks for them.

Element matrices are integrated exactly from 1-D Q1 integrals (tensor products), not by Gauss
points: for a box cell, ``D_de[A,B] = int dN_A/dx_d dN_B/dx_e`` factorises over the three axes, and
``K[(A,i),(B,j)] = sum_pq P[p,q] T[p,i,d] T[q,j,e] D_de[A,B]`` for a Voigt constitutive pattern P
(engineering shear order xx,yy,zz,xy,yz,xz, fe/solver.py:26-40).
"""
import functools

import numpy as np

LOC = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=np.int64)

# Voigt strain component p of displacement component i along derivative d
_T = np.zeros((6, 3, 3))
for _p, (_i, _d) in enumerate([(0, 0), (1, 1), (2, 2)]):
    _T[_p, _i, _d] = 1.0
for _p, (_a, _b) in zip((3, 4, 5), ((0, 1), (1, 2), (0, 2))):
    _T[_p, _a, _b] = 1.0
    _T[_p, _b, _a] = 1.0

P_LAMBDA = np.zeros((6, 6)); P_LAMBDA[:3, :3] = 1.0
P_MU = np.diag([2.0, 2.0, 2.0, 1.0, 1.0, 1.0])


def ortho_patterns():
    """z-orthotropic constitutive basis (fe/solver.py:43-52)."""
    P = np.zeros((6, 6, 6))
    P[0, 0, 0] = P[0, 1, 1] = 1
    P[1, 0, 1] = P[1, 1, 0] = 1
    P[2, [0, 2, 1, 2], [2, 0, 2, 1]] = 1
    P[3, 2, 2] = 1
    P[4, 3, 3] = 1
    P[5, 4, 4] = P[5, 5, 5] = 1
    return P


def _gradient_integrals(dims):
    """D[d, e, A, B] = int dN_A/dx_d dN_B/dx_e over the box (exact)."""
    mats = []
    for h in dims:
        m = h / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])          # int N_a N_b
        k = np.array([[1.0, -1.0], [-1.0, 1.0]]) / h               # int N_a' N_b'
        g = np.array([[-0.5, -0.5], [0.5, 0.5]])                   # int N_a' N_b
        mats.append((m, k, g))
    D = np.zeros((3, 3, 8, 8))
    for d in range(3):
        for e in range(3):
            fac = []
            for ax in range(3):
                m, k, g = mats[ax]
                if ax == d and ax == e:
                    fac.append(k)
                elif ax == d:
                    fac.append(g)
                elif ax == e:
                    fac.append(g.T)
                else:
                    fac.append(m)
            # node A = (a0, a1, a2) has index a0 + 2 a1 + 4 a2
            D[d, e] = np.kron(fac[2], np.kron(fac[1], fac[0]))
    return D


def element_from_pattern(dims, P):
    D = _gradient_integrals(np.asarray(dims, float))
    K = np.einsum("pq,pid,qje,deAB->AiBj", P, _T, _T, D)
    return K.reshape(24, 24)


@functools.lru_cache(maxsize=256)
def _family(dims_key):
    dims = np.array(dims_key)
    KL = element_from_pattern(dims, P_LAMBDA)
    KM = element_from_pattern(dims, P_MU)
    g = (2.0 * LOC - 1.0) / 4.0 / dims[None, :]                   # centre gradients (8,3)
    Bc = np.zeros((6, 24))
    for n in range(8):
        for p in range(6):
            for i in range(3):
                Bc[p, 3 * n + i] = sum(_T[p, i, d] * g[n, d] for d in range(3))
    KL_sri = dims.prod() * Bc.T @ P_LAMBDA @ Bc
    KB = np.stack([element_from_pattern(dims, P) for P in ortho_patterns()])
    return KL, KM, Bc, KL_sri, KB


def element_matrices(dims):
    """(K_lambda, K_mu, B_centre) of a box cell (fe/solver.py:77-90)."""
    KL, KM, Bc, _, _ = _family(tuple(np.round(np.asarray(dims, float), 12)))
    return KL, KM, Bc


def lame(E, nu):
    """fe/solver.py:131-134."""
    E = np.asarray(E, float)
    nu = np.asarray(nu, float)
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def ortho_coeffs(lam, mu, shear_factor=None):
    """fe/solver.py:55-60."""
    c = np.stack([lam + 2 * mu, lam, lam, lam + 2 * mu, mu, mu], 1)
    if shear_factor is not None:
        c[:, 5] *= shear_factor
    return c


class Grid:
    """Drop-in for fe/solver.py:93-128 (lazy heavy attributes)."""

    def __init__(self, xe, ye, ze):
        self.xe, self.ye, self.ze = (np.asarray(e, float) for e in (xe, ye, ze))
        self.nx, self.ny, self.nz = len(self.xe) - 1, len(self.ye) - 1, len(self.ze) - 1
        self.shape = (self.nx, self.ny, self.nz)
        self.ncell = self.nx * self.ny * self.nz
        self.nnode = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.dx, self.dy, self.dz = np.diff(self.xe), np.diff(self.ye), np.diff(self.ze)
        self.xc = (self.xe[:-1] + self.xe[1:]) / 2
        self.yc = (self.ye[:-1] + self.ye[1:]) / 2
        self.zc = (self.ze[:-1] + self.ze[1:]) / 2
        self._cache = {}

    def _lazy(self, key, fn):
        if key not in self._cache:
            self._cache[key] = fn()
        return self._cache[key]

    @property
    def cI(self):
        return self._lazy("cI", lambda: np.arange(self.ncell) // (self.ny * self.nz))

    @property
    def cJ(self):
        return self._lazy("cJ", lambda: (np.arange(self.ncell) // self.nz) % self.ny)

    @property
    def cK(self):
        return self._lazy("cK", lambda: np.arange(self.ncell) % self.nz)

    def nid(self, i, j, k):
        return i + (self.nx + 1) * (j + (self.ny + 1) * k)

    def cid(self, i, j, k):
        return (i * self.ny + j) * self.nz + k

    def corners(self, cells):
        """(n, 8) grid node ids of the given cells (local order a + 2b + 4c)."""
        cells = np.asarray(cells, np.int64)
        k = cells % self.nz
        j = (cells // self.nz) % self.ny
        i = cells // (self.ny * self.nz)
        base = self.nid(i, j, k)
        off = LOC[:, 0] + (self.nx + 1) * (LOC[:, 1] + (self.ny + 1) * LOC[:, 2])
        return base[:, None] + off[None, :]

    @property
    def cell_nodes(self):
        return self._lazy("cell_nodes", lambda: self.corners(np.arange(self.ncell)))

    def _dims(self):
        return np.stack([self.dx[self.cI], self.dy[self.cJ], self.dz[self.cK]], 1)

    def _families(self):
        def make():
            # unique dims families from the per-axis spacings (fe/solver.py:111-113)
            ux, ix = np.unique(np.round(self.dx, 6), return_inverse=True)
            uy, iy = np.unique(np.round(self.dy, 6), return_inverse=True)
            uz, iz = np.unique(np.round(self.dz, 6), return_inverse=True)
            code = (ix[self.cI] * len(uy) + iy[self.cJ]) * len(uz) + iz[self.cK]
            ucode, inv = np.unique(code, return_inverse=True)
            a, rem = np.divmod(ucode, len(uy) * len(uz))
            b, c = np.divmod(rem, len(uz))
            udims = np.stack([ux[a], uy[b], uz[c]], 1)
            return udims, inv.ravel()
        return self._lazy("fam", make)

    @property
    def udims(self):
        return self._families()[0]

    @property
    def dim_id(self):
        return self._families()[1]

    def _mats(self, idx):
        return np.stack([_family(tuple(np.round(d, 12)))[idx] for d in self.udims])

    @property
    def KL(self):
        return self._lazy("KL", lambda: self._mats(0))

    @property
    def KM(self):
        return self._lazy("KM", lambda: self._mats(1))

    @property
    def Bc(self):
        return self._lazy("Bc", lambda: self._mats(2))

    @property
    def KL_sri(self):
        return self._lazy("KLs", lambda: self._mats(3))

    @property
    def KB(self):
        return self._lazy("KB", lambda: self._mats(4))

    @property
    def cell_vol(self):
        return self._lazy("vol", lambda: self._dims().prod(1))

    @property
    def node_xyz(self):
        def make():
            n = np.arange(self.nnode)
            i = n % (self.nx + 1)
            j = (n // (self.nx + 1)) % (self.ny + 1)
            k = n // ((self.nx + 1) * (self.ny + 1))
            return np.stack([self.xe[i], self.ye[j], self.ze[k]], 1)
        return self._lazy("node_xyz", make)

    @property
    def cell_xyz(self):
        return self._lazy("cell_xyz", lambda: np.stack([self.xc[self.cI], self.yc[self.cJ], self.zc[self.cK]], 1))


def corners_of(grid, cells):
    """(n, 8) grid node ids of cells; works for this Grid and for the reference's fe/solver.Grid."""
    if hasattr(grid, "corners"):
        return grid.corners(cells)
    cells = np.asarray(cells, np.int64)
    k = cells % grid.nz
    j = (cells // grid.nz) % grid.ny
    i = cells // (grid.ny * grid.nz)
    base = i + (grid.nx + 1) * (j + (grid.ny + 1) * k)
    off = LOC[:, 0] + (grid.nx + 1) * (LOC[:, 1] + (grid.ny + 1) * LOC[:, 2])
    return base[:, None] + off[None, :]


def face_node_areas(grid, cells, face, cn=None):
    """Four face nodes per cell and tributary area face/4 (fe/solver.py:165-174,
    rb/splitfe.py:94-102 with ``cn``)."""
    ax = "xyz".index(face[0])
    side = 1 if face[1] == "+" else 0
    sel = LOC[:, ax] == side
    cells = np.asarray(cells, np.int64)
    nodes = (corners_of(grid, cells) if cn is None else cn[cells])[:, sel]
    k = cells % grid.nz
    j = (cells // grid.nz) % grid.ny
    i = cells // (grid.ny * grid.nz)
    d = [grid.dx[i], grid.dy[j], grid.dz[k]]
    o = [d[q] for q in range(3) if q != ax]
    return nodes.ravel(), np.repeat(o[0] * o[1] / 4.0, 4)
