"""Pins the FP64 oracle (CPU).  Every check is chosen so that a plausible mistake -- a dropped
term, a wrong sign, a swapped index, a transposed operand -- fails: closed forms, invariants,
brute force on tiny inputs and library routines (scipy splu), never a retyped formula."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import fe_oracle as O


def test_grid_numbering_closed_forms():
    """fe/solver.py:104-109,128: nodes x-fastest, cells z-fastest, corner order a + 2b + 4c."""
    g = O.Grid(np.arange(4.0), np.arange(3.0) * 2, np.arange(5.0))   # 3 x 2 x 4 cells
    assert g.nnode == 4 * 3 * 5 and g.ncell == 24
    c = g.cid(2, 1, 3)
    assert c == (2 * 2 + 1) * 4 + 3
    assert (g.cI[c], g.cJ[c], g.cK[c]) == (2, 1, 3)
    # corner (a,b,c) of cell (2,1,3) is node (2+a) + 4 ((1+b) + 3 (3+c))
    exp = [(2 + a) + 4 * ((1 + b) + 3 * (3 + cc)) for cc in (0, 1) for b in (0, 1) for a in (0, 1)]
    assert list(g.cell_nodes[c]) == exp
    assert np.allclose(g.node_xyz[exp[7]], [3.0, 4.0, 4.0])


@pytest.mark.parametrize("dims", [(1, 1, 1), (0.5, 0.7, 1.3), (2.0, 3.0, 0.25)])
def test_element_matrices_invariants(dims):
    KL, KM, Bc = O.element_matrices(dims)
    for K in (KL, KM):
        assert np.allclose(K, K.T, atol=1e-13)
    # six rigid-body modes are in the null space of lam KL + mu KM
    X = np.array([[a * dims[0], b * dims[1], c * dims[2]] for c in (0, 1) for b in (0, 1) for a in (0, 1)])
    R = []
    for d in range(3):
        v = np.zeros((8, 3)); v[:, d] = 1; R.append(v.ravel())
    for (i, j) in ((0, 1), (1, 2), (0, 2)):
        v = np.zeros((8, 3)); v[:, i] = -X[:, j]; v[:, j] = X[:, i]; R.append(v.ravel())
    K = 1.3 * KL + 0.7 * KM
    for v in R:
        assert np.abs(K @ v).max() < 1e-12 * np.abs(K).max()
    assert np.linalg.matrix_rank(K, tol=1e-10 * np.abs(K).max()) == 18
    # energy of a linear displacement u = G x equals V * eps:D:eps (closed form, exact integration)
    rng = np.random.default_rng(0)
    G = rng.standard_normal((3, 3))
    u = (X @ G.T).ravel()
    eps = 0.5 * (G + G.T)
    V = np.prod(dims)
    lam, mu = 1.3, 0.7
    e_exact = V * (lam * np.trace(eps) ** 2 + 2 * mu * (eps ** 2).sum())
    assert np.isclose(u @ K @ u, e_exact, rtol=1e-12)
    # centre B gives the strain of the linear field in engineering order
    e6 = Bc @ u
    assert np.allclose(e6, [G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])


def test_kl_sri_rank_one():
    dims = (0.5, 0.7, 1.3)
    K = O.kl_sri(dims)
    assert np.linalg.matrix_rank(K, tol=1e-10 * np.abs(K).max()) == 1
    Bc = O.element_matrices(dims)[2]
    g = Bc[:3].sum(0)                      # divergence at the centre
    assert np.allclose(K, np.prod(dims) * np.outer(g, g))


def test_assemble_brute_force():
    g = O.Grid([0, 1, 2.5], [0, 1], [0, 0.5, 1.5])
    rng = np.random.default_rng(1)
    lam, mu = rng.uniform(1, 2, g.ncell), rng.uniform(1, 2, g.ncell)
    active = np.array([True, True, False, True])
    K, nmap, used, cells, _ = O.assemble(g, lam, mu, active)
    D = np.zeros((3 * g.nnode, 3 * g.nnode))
    for c in np.flatnonzero(active):
        d = g.dim_id[c]
        dofs = (3 * g.cell_nodes[c][:, None] + np.arange(3)).ravel()
        D[np.ix_(dofs, dofs)] += lam[c] * g.KL[d] + mu[c] * g.KM[d]
    keep = (3 * used[:, None] + np.arange(3)).ravel()
    assert np.allclose(K.toarray(), D[np.ix_(keep, keep)], atol=1e-13)


def _random_bones(seed, shape=(6, 5, 4), nb=3):
    rng = np.random.default_rng(seed)
    g = O.Grid(*(np.arange(n + 1.0) for n in shape))
    part = rng.integers(0, nb + 1, g.ncell)
    active = rng.uniform(size=g.ncell) > 0.15
    share = rng.uniform(size=(g.ncell, nb + 1)).astype(np.float32)
    share[rng.uniform(size=g.ncell) < 0.3] = 0.0                 # cells holding none of the bones
    share[:, 0] = 0
    return g, active, part, share


@pytest.mark.parametrize("seed", range(6))
def test_split_connectivity_fast_equals_loop(seed):
    """rb/splitfe.py:262 says the vectorised form gives the same result as the loop form."""
    g, active, part, share = _random_bones(seed)
    a = O.split_connectivity(g, active, part, share)
    b = O.split_connectivity_fast(g, active, part, share)
    assert np.array_equal(a[0][active], b[0][active]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_split_connectivity_two_bones_by_hand():
    """Two cells of different bones share a face: its 4 nodes are split, the second bone gets
    copies nnode + 0..3 in node order; a part-0 cell attaches to the bone it holds most of."""
    g = O.Grid([0, 1, 2, 3], [0, 1], [0, 1])          # cells 0,1,2 along x
    part = np.array([1, 2, 0])
    share = np.zeros((3, 3), np.float32)
    share[2, 1], share[2, 2] = 0.2, 0.6
    cn, orig, ns = O.split_connectivity(g, np.ones(3, bool), part, share)
    face = [1, 5, 9, 13]                               # x = 1 nodes
    assert ns == 4 and list(orig[g.nnode:]) == face
    assert list(cn[1][[0, 2, 4, 6]]) == [g.nnode + q for q in range(4)]
    assert list(cn[0]) == list(g.cell_nodes[0])
    # cell 2 (part 0) touches x=2 nodes only: none split, untouched
    assert list(cn[2]) == list(g.cell_nodes[2])


def test_p1d_linear_and_partition_of_unity():
    x = np.array([0.0, 0.3, 1.0, 1.1, 2.5, 2.6])
    P, xc = O.p1d(x)
    assert list(xc) == [0.0, 1.0, 2.5, 2.6]            # every other node plus the last
    assert np.allclose(P.sum(1), 1.0)
    assert np.allclose(P @ (3 * xc - 1), 3 * x - 1)


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return sp.csr_matrix(A @ A.T + n * np.eye(n))


def test_chebyshev_smoother_polynomial():
    """x_new = x + (I - p(MA)) (x* - x), p(t) = T_3((theta - t)/delta) / T_3(theta/delta)."""
    A = _spd(12, 2)
    L = O.Level(A)
    rng = np.random.default_rng(3)
    b, x = rng.standard_normal(12), rng.standard_normal(12)
    xs = np.linalg.solve(A.toarray(), b)
    Minv = np.zeros((12, 12))
    for i in range(4):
        Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = L.Dinv[i]
    MA = Minv @ A.toarray()
    th, de = 0.5 * (L.lmax + L.lmin), 0.5 * (L.lmax - L.lmin)
    T3 = lambda s: 4 * s ** 3 - 3 * s
    w, V = np.linalg.eig(MA)
    P = (V @ np.diag(T3((th - w) / de) / T3(th / de)) @ np.linalg.inv(V)).real
    assert np.allclose(xs - L.smooth(x, b), P @ (xs - x), atol=1e-10)


def test_block_jacobi_sweeps():
    A = _spd(9, 4)
    L = O.Level(A, smoother="jacobi", omega=0.6, sweeps=2)
    b, x = np.ones(9), np.zeros(9)
    Minv = np.zeros((9, 9))
    for i in range(3):
        Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.linalg.inv(A.toarray()[3 * i:3 * i + 3, 3 * i:3 * i + 3])
    E = np.eye(9) - 0.6 * Minv @ A.toarray()
    xs = np.linalg.solve(A.toarray(), b)
    assert np.allclose(xs - L.smooth(x, b), E @ E @ (xs - x))


def test_two_grid_error_propagation():
    """V-cycle on 2 levels: E = S (I - P Ac^-1 P^T A) S, with S the smoother error operator."""
    g = O.Grid(np.arange(5.0), np.arange(3.0), np.arange(3.0))
    lam, mu = np.full(g.ncell, 1.0), np.full(g.ncell, 1.0)
    K, nmap, used, _, _ = O.assemble(g, lam, mu, np.ones(g.ncell, bool))
    K = K + sp.identity(K.shape[0]) * 0.5
    mg = O.GMG(K, used, g.xe, g.ye, g.ze, levels=2, smoother="jacobi", omega=0.6, sweeps=1)
    n = K.shape[0]
    A = K.toarray()
    Minv = np.zeros((n, n))
    for i in range(n // 3):
        Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = mg.levels[0].Dinv[i]
    S = np.eye(n) - 0.6 * Minv @ A
    P = mg.P[0].toarray()
    E = S @ (np.eye(n) - P @ np.linalg.solve(P.T @ A @ P, P.T @ A)) @ S
    rng = np.random.default_rng(5)
    e = rng.standard_normal(n)
    b = A @ e
    # vcycle(b) from x=0 approximates e: error e - vcycle(b) = E e
    assert np.allclose(e - mg.vcycle(b), E @ e, atol=1e-10)


def test_pcg_gmg_matches_splu_and_cadence():
    g = O.Grid(np.arange(9.0), np.arange(7.0), np.arange(6.0))
    rng = np.random.default_rng(6)
    E = np.where(rng.uniform(size=g.ncell) < 0.3, 17000.0, 700.0)
    lam, mu = O.lame(E, np.full(g.ncell, 0.3))
    bot = np.flatnonzero(g.cK == 0); top = np.flatnonzero(g.cK == g.nz - 1)
    nb, ab = O.face_node_areas(g, bot, "z-"); nt, at = O.face_node_areas(g, top, "z+")
    args = (g, lam, mu, np.ones(g.ncell, bool), [(nb, ab, np.array([1e6] * 3))], [(nt, at, np.array([0, 0, -1.0]))])
    Ud, _ = O.solve(*args, method="direct")
    Ug, info = O.solve(*args, method="gmg", tol=1e-12)
    assert np.linalg.norm(Ug - Ud) / np.linalg.norm(Ud) < 1e-9 and info["flag"] == 0
    _, inf0 = O.solve(*args, method="none")
    K, f = inf0["K"], inf0["f"]
    mg = O.GMG(K, inf0["used"], g.xe, g.ye, g.ze)
    x, it, flag, rr = O.pcg_cadence(K, f, mg.vcycle, 1e-10, 500)
    assert flag == 0 and rr <= 1e-10 and (it < 5 or it % 5 == 0)
    assert np.linalg.norm(K @ x - f) <= 1e-10 * np.linalg.norm(f)


def test_patch_test_closed_forms():
    """fe/test_patch.py:4-14."""
    g = O.Grid(np.r_[0, 1, 3, 4, 7, 10], np.r_[0, 2, 5, 10], np.r_[0, 1, 2, 4, 6, 10])
    E, nu = 1000.0, 0.3
    lam, mu = O.lame(np.full(g.ncell, E), np.full(g.ncell, nu))
    bot, top = np.flatnonzero(g.cK == 0), np.flatnonzero(g.cK == g.nz - 1)
    nb, ab = O.face_node_areas(g, bot, "z-")
    nt, at = O.face_node_areas(g, top, "z+")
    U, info = O.solve(g, lam, mu, np.ones(g.ncell, bool), [(nb, ab, np.array([1e-6, 1e-6, 1e9]))],
                      [(nt, at, np.array([0, 0, -1.0]))], method="direct")
    sig, vm, eps = O.cell_stress(g, lam, mu, U, np.arange(g.ncell))
    top_nodes = np.flatnonzero(g.node_xyz[:, 2] == 10)
    assert np.allclose(U[top_nodes, 2], -10 / E, rtol=1e-6)
    assert np.allclose(sig[:, 2], -1.0, atol=1e-8) and np.allclose(vm, 1.0, atol=1e-8)
    assert abs(np.ptp(U[:, 0]) - nu * 10 / E) < 1e-9


def _ortho_D(c):
    """Stiffness of a material symmetric about the z axis in Voigt order (xx, yy, zz, xy, yz, xz),
    written as the matrix: c0 = D11 = D22, c1 = D12, c2 = D13 = D23, c3 = D33, c4 = D44 (xy),
    c5 = D55 = D66 (yz, xz)."""
    c0, c1, c2, c3, c4, c5 = c
    return np.array([[c0, c1, c2, 0, 0, 0], [c1, c0, c2, 0, 0, 0], [c2, c2, c3, 0, 0, 0],
                     [0, 0, 0, c4, 0, 0], [0, 0, 0, 0, c5, 0], [0, 0, 0, 0, 0, c5]], float)


def test_orthotropic_affine_energy_and_stress():
    """fe/solver.py:43-74: for u = G x (any displacement gradient G) the trilinear element holds the
    field exactly and 2x2x2 Gauss integrates its constant strain exactly, so u^T (sum_k c_k KB_k) u =
    V eps^T D eps and the centre stress is D eps; eps in engineering shear (xy, yz, xz)."""
    rng = np.random.default_rng(3)
    g = O.Grid(np.r_[0, 1.5, 2.0], np.r_[0, 0.7], np.r_[0, 2.5, 3.0])
    c = np.array([9.0, 2.0, 1.5, 6.0, 3.0, 0.8])
    coeffs = np.tile(c, (g.ncell, 1))
    K, nmap, used, cells, dofs = O.assemble(g, None, None, np.ones(g.ncell, bool), coeffs)
    D = _ortho_D(c)
    assert np.allclose(np.einsum("k,kij->ij", c, O.ortho_patterns()), D)
    for _ in range(4):
        G = rng.standard_normal((3, 3))
        U = g.node_xyz @ G.T
        u = U[used].ravel()
        eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0]])
        vol = g.cell_vol.sum()
        assert np.isclose(u @ K @ u, vol * eps @ D @ eps, rtol=1e-12)
        sig, vm, e = O.cell_stress(g, None, None, U, np.arange(g.ncell), coeffs=coeffs)
        assert np.allclose(e, eps[None], atol=1e-12) and np.allclose(sig, (D @ eps)[None], rtol=1e-12, atol=1e-12)
    iso = O.ortho_coeffs(np.full(g.ncell, 2.0), np.full(g.ncell, 3.0))
    Ki, *_ = O.assemble(g, None, None, np.ones(g.ncell, bool), iso)
    Kl, *_ = O.assemble(g, np.full(g.ncell, 2.0), np.full(g.ncell, 3.0), np.ones(g.ncell, bool))
    assert abs(Ki - Kl).max() <= 1e-12 * abs(Kl).max()


def test_orthotropic_column_compliance():
    """Uniaxial z load on a free-standing column of a z-symmetric material: uniform stress
    (0, 0, -1, 0, 0, 0), strain S sigma with S = D^-1, top settlement height * S33, lateral
    spread width * S13."""
    g = O.Grid(np.r_[0, 1, 3, 4], np.r_[0, 2, 5], np.r_[0, 1, 2, 4, 6, 10])
    c = np.array([900.0, 300.0, 250.0, 600.0, 200.0, 60.0])
    coeffs = np.tile(c, (g.ncell, 1))
    bot, top = np.flatnonzero(g.cK == 0), np.flatnonzero(g.cK == g.nz - 1)
    nb, ab = O.face_node_areas(g, bot, "z-")
    nt, at = O.face_node_areas(g, top, "z+")
    pin = np.flatnonzero(np.all(g.node_xyz == 0, 1))                # x, y of the origin and y of
    pin_x = np.flatnonzero(np.all(g.node_xyz == [4, 0, 0], 1))      # (4, 0, 0): no rigid motion left
    U, info = O.solve(g, None, None, np.ones(g.ncell, bool),
                      [(nb, ab, np.array([0.0, 0.0, 1e12])), (pin, np.ones(1), np.array([1e6, 1e6, 0.0])),
                       (pin_x, np.ones(1), np.array([0.0, 1e6, 0.0]))],
                      [(nt, at, np.array([0, 0, -1.0]))], method="direct", coeffs=coeffs)
    S = np.linalg.inv(_ortho_D(c))
    sig, vm, eps = O.cell_stress(g, None, None, U, np.arange(g.ncell), coeffs=coeffs)
    assert np.allclose(sig, [[0, 0, -1.0, 0, 0, 0]], atol=1e-8)
    assert np.allclose(eps, (S @ [0, 0, -1.0, 0, 0, 0])[None], rtol=1e-8, atol=1e-14)
    top_nodes = np.flatnonzero(g.node_xyz[:, 2] == 10)
    assert np.allclose(U[top_nodes, 2], -10 * S[2, 2], rtol=1e-8)
    assert np.isclose(np.ptp(U[:, 0]), 4 * abs(S[0, 2]), rtol=1e-8)


def test_von_mises_closed_forms():
    s = np.array([[-1, 0, 0, 0, 0, 0], [5, 5, 5, 0, 0, 0], [0, 0, 0, 2, 0, 0], [1, 2, 3, 0, 0, 0.0]])
    vm = O.von_mises(s)
    assert np.allclose(vm[:3], [1.0, 0.0, 2 * np.sqrt(3)])
    assert np.isclose(vm[3], np.sqrt(3.0))


def test_spring_matrix_closed_form():
    """One ligament between single-node patches: K = k w w^T with w = +d at a, -d at b."""
    xyz = np.array([[0, 0, 0], [10.0, 0, 0], [0, 20, 0], [20, 20, 20.0]])
    springs = [dict(a=[0, 0, 0.0], b=[3.0, 4.0, 0], k=5.0, A="x", B="y")]
    build = O.spring_matrix_cn(xyz, springs, {"x": np.array([0]), "y": np.array([2])})
    nmap = np.arange(4)
    K = build(nmap, 12).toarray()
    d = np.array([0.6, 0.8, 0.0])
    w = np.zeros(12); w[0:3] = d; w[6:9] = -d
    assert np.allclose(K, 5.0 * np.outer(w, w))


def test_sample_fe_constant_and_linear():
    g = O.Grid(np.arange(5.0), np.arange(5.0), np.arange(5.0))
    gm = O.Grid(np.arange(0, 5, 2.0), np.arange(0, 5, 2.0), np.arange(0, 5, 2.0))
    active = np.ones(g.ncell, bool)
    VM = np.full(g.ncell, 3.0, np.float32)
    SIG = np.tile(np.arange(6, dtype=np.float32), (g.ncell, 1))
    Ur = g.node_xyz @ np.array([[1, 0, 0], [0, 2, 0], [0, 0, 3.0]])      # linear field
    vm, u, cov, sig = O.sample_fe(g, g.cell_nodes, active, VM, SIG, Ur, gm)
    assert np.allclose(vm, 3.0) and np.allclose(cov, 1.0) and np.allclose(sig, np.arange(6))
    centres = gm.cell_xyz * np.array([1, 2, 3.0])
    assert np.allclose(u, centres, atol=1e-6)


@pytest.mark.parametrize("density", [0.03, 0.25])
def test_edt_feature_transform_is_nearest_with_kji_ties(density):
    """The nearest-active-cell map sample_nodes takes from scipy's feature transform equals
    enumeration: smallest squared index distance, ties to the smallest (k, j, i)."""
    from scipy import ndimage
    rng = np.random.default_rng(int(density * 100))
    for shape in ((7, 5, 6), (9, 4, 8), (5, 8, 7)):
        act = rng.uniform(size=shape) < density
        act[0, 0, 0] = True
        dist, feat = ndimage.distance_transform_edt(~act, return_indices=True)
        cells = np.arange(act.size)
        best, d2 = O.nearest_active_bruteforce(act.ravel(), shape, cells)
        assert np.array_equal(feat.reshape(3, -1).T, best)
        assert np.allclose(dist.ravel() ** 2, d2)


def test_edt_tie_rule_against_stored_scipy_indices():
    """tests/golden/edt_ties.npz holds scipy's feature transform of three tie-rich rasters: the
    enumeration rule (smallest squared distance, then smallest (k, j, i)) reproduces it, and the
    installed scipy still does."""
    import os
    from scipy import ndimage
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "edt_ties.npz"))
    for r in range(3):
        act, ind = z[f"active{r}"], z[f"indices{r}"].astype(np.int64)
        best, _ = O.nearest_active_bruteforce(act.ravel(), act.shape, np.arange(act.size))
        assert np.array_equal(best, ind.reshape(3, -1).T)
        _, now = ndimage.distance_transform_edt(~act, return_indices=True)
        assert np.array_equal(now, ind)


def test_sample_nodes_nearest_cell_and_linear_field():
    """Nodes in active cells reproduce a linear field; nodes in empty cells take the field of the
    brute-force nearest active cell at the clamped local coordinates."""
    rng = np.random.default_rng(5)
    g = O.Grid(np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 8)]), np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 6)]),
               np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 7)]))
    act = rng.uniform(size=g.ncell) < 0.2
    act[0] = True
    B = np.array([[1.0, 0.5, -0.25], [0.3, 2.0, 0.1], [-0.7, 0.2, 3.0]])
    Ur = g.node_xyz @ B.T + np.array([0.1, -0.2, 0.3])
    gm = O.Grid(np.linspace(-1.0, g.xe[-1] + 1.0, 6), np.linspace(-0.5, g.ye[-1] + 0.7, 5),
                np.linspace(0.2, g.ze[-1] + 1.5, 7))
    u = O.sample_nodes(g, g.cell_nodes, act, Ur, gm)
    P = gm.node_xyz
    held = [np.clip(np.searchsorted(e, P[:, a]) - 1, 0, n - 1)
            for a, (e, n) in enumerate(((g.xe, g.nx), (g.ye, g.ny), (g.ze, g.nz)))]
    cells = (held[0] * g.ny + held[1]) * g.nz + held[2]
    near, _ = O.nearest_active_bruteforce(act, g.shape, cells)
    lo = np.stack([g.xe[near[:, 0]], g.ye[near[:, 1]], g.ze[near[:, 2]]], 1)
    hi = np.stack([g.xe[near[:, 0] + 1], g.ye[near[:, 1] + 1], g.ze[near[:, 2] + 1]], 1)
    Q = np.clip(P, lo, hi)                        # clamped point inside the chosen cell
    expect = Q @ B.T + np.array([0.1, -0.2, 0.3])
    assert np.abs(u - expect).max() <= 1e-5 * np.abs(expect).max()
    inside = act[cells] & (P == Q).all(1)
    assert inside.sum() > 0 and (~act[cells]).sum() > 20


# ---- lattice materials (rb/rbfe.py:75-134) ---------------------------------------------------
LAT_CLS = {1: 1, 2: 2, 3: 1, 4: 2, 5: 3, 6: 4, 7: 5, 8: 6, 9: 7, 10: 8, 11: 9, 12: 1}
LAT_ROLES = {"bones": {"calc": (1, 2), "talus": (3, 4)}, "bonded_extra": [12]}


def _lat_E_nu(name, cart_E=10.0, nu_soft=0.49):
    if name == "cartilage":
        return cart_E, 0.4
    if name == "soft":
        return 1.15, nu_soft
    return O.LATTICE_MAT[name]


def _lattice(dims, fill):
    nc = int(np.prod(dims))
    return np.full((nc, 27), fill, np.int64)


def test_lattice_single_phase_cells_return_the_phase():
    """A cell of one class gets that class's (E, nu); void, gap and plate cells (plate is not one of
    the occupied phases of rb/rbfe.py:96-99) are inactive (E=1, nu=0.3)."""
    dims = (3, 1, 1)
    for lab, cname in [(1, "cortical"), (2, "cancellous"), (5, "cartilage"), (6, "fascia"), (7, "soft"),
                       (9, "fibrous"), (10, "fat")]:
        L = _lattice(dims, lab)
        E, nu, act, F, ser = O.lattice_materials(L, LAT_CLS, LAT_ROLES, dims, cart_E=12.5, nu_soft=0.47)
        Ek, nk = _lat_E_nu(cname, 12.5, 0.47)
        assert act.all(), cname
        assert np.allclose(E, Ek, rtol=1e-12) and np.allclose(nu, nk, rtol=1e-12), (cname, E, nu)
    for lab in (0, 11, 8):                                # void, gap, plate
        L = _lattice(dims, lab)
        E, nu, act, F, ser = O.lattice_materials(L, LAT_CLS, LAT_ROLES, dims)
        assert not act.any() and np.all(E == 1.0) and np.all(nu == 0.3)


def test_lattice_hard_in_series_with_cartilage_is_reuss():
    """Cortical fraction h and cartilage 1-h in one cell: constrained modulus M and shear G are the
    series (Reuss) averages of the two phases."""
    dims = (1, 1, 1)
    L = _lattice(dims, 1)
    L[0, 18:] = 5                                         # 9 of 27 samples cartilage
    E, nu, act, F, ser = O.lattice_materials(L, LAT_CLS, LAT_ROLES, dims, cart_E=10.0)
    h = 18 / 27
    lam = lambda E_, n_: E_ * n_ / ((1 + n_) * (1 - 2 * n_))
    mu = lambda E_, n_: E_ / (2 * (1 + n_))
    Ec, nc = O.LATTICE_MAT["cortical"]
    Mh, Gh = lam(Ec, nc) + 2 * mu(Ec, nc), mu(Ec, nc)
    Ml, Gl = lam(10.0, 0.4) + 2 * mu(10.0, 0.4), mu(10.0, 0.4)
    M = 1 / (h / Mh + (1 - h) / Ml)
    G = 1 / (h / Gh + (1 - h) / Gl)
    l_, m_ = M - 2 * G, G
    assert ser[0] and act[0]
    assert np.isclose(E[0], m_ * (3 * l_ + 2 * m_) / (l_ + m_), rtol=1e-12)
    assert np.isclose(nu[0], l_ / (2 * (l_ + m_)), rtol=1e-12)


def test_lattice_two_phases_in_a_plain_cell_is_voigt():
    """Cortical and soft tissue in a cell that is no joint: Lame parameters are the volume (Voigt)
    averages."""
    dims = (1, 1, 1)
    L = _lattice(dims, 1)
    L[0, 20:] = 7
    E, nu, act, F, ser = O.lattice_materials(L, LAT_CLS, LAT_ROLES, dims, nu_soft=0.49)
    h = 20 / 27
    lam = lambda E_, n_: E_ * n_ / ((1 + n_) * (1 - 2 * n_))
    mu = lambda E_, n_: E_ / (2 * (1 + n_))
    l_ = h * lam(*O.LATTICE_MAT["cortical"]) + (1 - h) * lam(1.15, 0.49)
    m_ = h * mu(*O.LATTICE_MAT["cortical"]) + (1 - h) * mu(1.15, 0.49)
    assert not ser[0] and act[0]
    assert np.isclose(E[0], m_ * (3 * l_ + 2 * m_) / (l_ + m_), rtol=1e-12)
    assert np.isclose(nu[0], l_ / (2 * (l_ + m_)), rtol=1e-12)


def test_lattice_joint_cells_and_activity_against_brute_force():
    """ser (joint) = cartilage present, or >= 2 bone groups within the 3x3x3 neighbourhood (brute
    force); active = occupancy > 0.15 in joint cells, >= 0.5 elsewhere."""
    rng = np.random.default_rng(4)
    dims = (7, 6, 5)
    nc = int(np.prod(dims))
    L = rng.choice([0, 0, 0, 1, 2, 3, 4, 7, 10, 12, 13], size=(nc, 27))
    cls = dict(LAT_CLS)
    cls[13] = 2                                            # an unnamed bone body: its own group
    E, nu, act, F, ser = O.lattice_materials(L, cls, LAT_ROLES, dims)
    grp, names = O.bone_groups(cls, LAT_ROLES)
    assert grp == {1: 1, 2: 1, 3: 2, 4: 2, 13: 3} and names[3] == "body13"
    gl = np.zeros(14, np.int64)
    for b, g in grp.items():
        gl[b] = g
    pres = [set(gl[L[c]][gl[L[c]] > 0]) for c in range(nc)]
    nx, ny, nz = dims
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                c = (i * ny + j) * nz + k
                near = set()
                for di in (-1, 0, 1):
                    for dj in (-1, 0, 1):
                        for dk in (-1, 0, 1):
                            a, b, d = i + di, j + dj, k + dk
                            if 0 <= a < nx and 0 <= b < ny and 0 <= d < nz:
                                near |= pres[(a * ny + b) * nz + d]
                cart = (L[c] == 5).any() or (L[c] == 9).any()
                assert ser[c] == (cart or len(near) >= 2), (i, j, k)
                occ = np.isin(L[c], [1, 2, 3, 4, 6, 5, 9, 7, 10, 12, 13]).sum() / 27
                assert act[c] == ((occ > 0.15) if ser[c] else (occ >= 0.5)), (i, j, k)


def test_lattice_bone_share_and_part():
    """Shares are the per-group sample fractions; part = the group holding >= 50 % of a cell without
    joint layers (rb/replica3.py:17-29, rb/rbfe.py:140-141)."""
    L = np.array([[1] * 14 + [3] * 13, [3] * 20 + [5] * 7, [7] * 27, [1] * 10 + [3] * 10 + [7] * 7])
    S, names = O.bone_share(L, LAT_CLS, LAT_ROLES)
    assert S.dtype == np.float32 and S.shape == (4, 3)
    assert np.allclose(S[0], [0, 14 / 27, 13 / 27]) and np.allclose(S[3], [0, 10 / 27, 10 / 27])
    F = np.stack([(np.vectorize(lambda b: LAT_CLS.get(b, 0))(L) == k).mean(1) for k in range(10)], 1)
    part = O.lattice_part(S, F)
    assert list(part) == [1, 0, 0, 0]                     # row 1: cartilage present; rows 2-3: < 50 %


def test_prolongation_axis_mask():
    """A kept axis (mask bit clear) has the identity prolongation: with every axis kept P is the
    identity on the active DOFs; with x kept only, P equals the reference prolongation of the
    other two axes kron the identity in x; linear fields are reproduced exactly either way."""
    xe, ye, ze = np.r_[0.0, 1, 3, 4, 7], np.r_[0.0, 2, 5], np.r_[0.0, 1, 2, 4]
    g = O.Grid(xe, ye, ze)
    act = np.arange(3 * g.nnode)
    P0, keep0, c0 = O.prolongation((xe, ye, ze), act, axes=0)
    assert (P0 - sp.identity(3 * g.nnode)).count_nonzero() == 0 and len(keep0) == 3 * g.nnode
    P6, keep6, c6 = O.prolongation((xe, ye, ze), act, axes=6)
    Py, _ = O.p1d(ye)
    Pz, _ = O.p1d(ze)
    ref = sp.kron(sp.kron(Pz, sp.kron(Py, sp.identity(len(xe)))), sp.identity(3)).tocsr()
    assert abs(P6 - ref[:, keep6]).max() < 1e-15
    assert np.array_equal(c6[0], xe) and len(c6[1]) < len(ye)
    # linear field a + b x + c y + d z in component 1
    X, Y, Z = np.meshgrid(c6[0], c6[1], c6[2], indexing="ij")
    fc = (1.0 + 2 * X - 3 * Y + 0.5 * Z).transpose(2, 1, 0).ravel()
    uc = np.zeros((len(fc), 3)); uc[:, 1] = fc
    Xf, Yf, Zf = np.meshgrid(xe, ye, ze, indexing="ij")
    ff = (1.0 + 2 * Xf - 3 * Yf + 0.5 * Zf).transpose(2, 1, 0).ravel()
    uf = (P6 @ uc.ravel()[keep6]).reshape(-1, 3)
    assert np.abs(uf[:, 1] - ff).max() < 1e-12 and not uf[:, 0].any()


def _engine_grid(seed=0, shape=(7, 6, 5)):
    rng = np.random.default_rng(seed)
    xe, ye, ze = (np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, n)]) for n in shape)
    return O.Grid(xe, ye, ze), rng


def test_bone_node_sets_by_enumeration():
    g, rng = _engine_grid(1)
    part = rng.integers(0, 4, g.ncell)
    active = rng.uniform(size=g.ncell) > 0.2
    sets = O.bone_node_sets(g.cell_nodes, part, active, [1, 2, 3])
    counts = O.part_cell_counts(part, active, [1, 2, 3])
    for k in (1, 2, 3):
        ref = set()
        for c in range(g.ncell):
            if part[c] == k and active[c]:
                ref.update(int(n) for n in g.cell_nodes[c])
        assert list(sets[k]) == sorted(ref)
        assert counts[k] == sum(1 for c in range(g.ncell) if part[c] == k and active[c])


def test_ground_contact_area_closed_form():
    g, rng = _engine_grid(2)
    active = rng.uniform(size=g.ncell) > 0.3
    contact = rng.uniform(size=(g.nx, g.ny)) * (rng.uniform(size=(g.nx, g.ny)) > 0.4)
    nodes, area = O.ground_contact(g, g.cell_nodes, active, contact)
    ref = sum(g.dx[i] * g.dy[j] * contact[i, j] for i in range(g.nx) for j in range(g.ny)
              if contact[i, j] > 0 and active[g.cid(i, j, 0)])
    assert abs(area.sum() - ref) < 1e-12 * max(ref, 1)
    assert np.all(g.node_xyz[nodes][:, 2] == g.ze[0])


@pytest.mark.parametrize("through_centroid", [False, True])
def test_leg_loads_resultant_on_load_line(through_centroid):
    """The pressure's resultant is -Fz (times load) and passes through the load line; when the line
    passes through the loaded area's centroid the pressure is uniform."""
    g, rng = _engine_grid(3)
    active = np.ones(g.ncell, bool)
    frac = {"tibia": rng.uniform(size=g.ncell), "fibula": rng.uniform(size=g.ncell)}
    top = np.flatnonzero(g.cK == g.nz - 1)
    lines = {}
    for nm in frac:
        c = top[frac[nm][top] > 0.5]
        n_, a_ = O.face_node_areas(g, c, "z+")
        X = g.node_xyz[n_][:, :2]
        xb = (a_[:, None] * X).sum(0) / a_.sum()
        lines[nm] = xb if through_centroid else xb + rng.uniform(-0.3, 0.3, 2)
    forces = {"tibia": 300.0, "fibula": 60.0}
    out = O.leg_loads(g, g.cell_nodes, np.arange(g.nnode), active, frac, forces, 1.5, lines)
    for (n_, f, d), nm in zip(out, forces):
        assert np.array_equal(d, [0, 0, 1])
        assert abs(f.sum() + 1.5 * forces[nm]) < 1e-9 * forces[nm]
        X = g.node_xyz[n_][:, :2]
        M = (f[:, None] * X).sum(0)
        assert np.abs(M + 1.5 * forces[nm] * np.asarray(lines[nm])).max() < 1e-9 * forces[nm] * 10
        if through_centroid:
            _, a_ = O.face_node_areas(g, top[frac[nm][top] > 0.5], "z+")
            assert np.allclose(f / a_, f[0] / a_[0], rtol=1e-9)


def test_spring_end_bones_match_kdtree():
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(4)
    xyz = rng.uniform(0, 50, (4000, 3))
    bone_nodes = {"a": np.sort(rng.choice(4000, 700, replace=False)), "b": np.sort(rng.choice(4000, 900, replace=False)),
                  "empty": np.zeros(0, np.int64), "c": np.sort(rng.choice(4000, 300, replace=False))}
    springs = [{"a": rng.uniform(0, 50, 3).tolist(), "b": rng.uniform(0, 50, 3).tolist(), "k": 1.0} for _ in range(40)]
    # an end exactly on a node shared by two bones: the first bone in order wins the tie
    shared = np.intersect1d(bone_nodes["a"], bone_nodes["b"])[0]
    springs.append({"a": xyz[shared].tolist(), "b": xyz[bone_nodes["c"][0]].tolist(), "k": 1.0})
    got = O.spring_end_bones(springs, bone_nodes, xyz)
    trees = {nm: cKDTree(xyz[nd]) for nm, nd in bone_nodes.items() if len(nd)}
    for s, (ea, eb) in zip(springs, got):
        assert ea == min(trees, key=lambda nm: trees[nm].query(s["a"])[0])
        assert eb == min(trees, key=lambda nm: trees[nm].query(s["b"])[0])
    assert got[-1] == ("a", "c")


def _free_body(seed=0, dims=(4, 3, 5)):
    # uneven spacings exact at 6 decimals: the element families of fe/solver.py:112 are built for
    # the cell sizes rounded to 6 decimals, so other sizes make rigid motions cost ~1e-9 of |K|
    rng = np.random.default_rng(seed)
    sizes = np.array([0.5, 0.75, 1.0, 1.25, 2.0])
    g = O.Grid(*[np.cumsum(np.r_[0, rng.choice(sizes, n)]) for n in dims])
    lam, mu = O.lame(rng.uniform(100, 2000, g.ncell), rng.uniform(0.1, 0.45, g.ncell))
    return g, lam, mu


def test_rigid_modes_span_the_kernel_of_a_free_body():
    """fe/solver.py:177-199: on a body without springs the 3 translations and 3 rotations of a
    group are exact zero-energy modes, K Z = 0 (a wrong sign, a swapped component or an
    uncentred/misplaced rotation axis gives K Z != 0), and the 6 columns are independent."""
    g, lam, mu = _free_body()
    K, nmap, used, _, _ = O.assemble(g, lam, mu, np.ones(g.ncell, bool))
    Z = O.rigid_modes(g.node_xyz, nmap, K.shape[0], [np.arange(g.nnode)]).toarray()
    assert Z.shape == (K.shape[0], 6)
    assert np.abs(K @ Z).max() <= 1e-11 * abs(K).max() * np.abs(Z).max()
    assert np.linalg.matrix_rank(Z) == 6
    # a genuine deformation (stretch along x) is not in the span of the modes
    x = g.node_xyz[used]
    stretch = np.zeros((len(used), 3)); stretch[:, 0] = x[:, 0]
    assert np.abs(K @ stretch.ravel()).max() > 1e-3 * abs(K).max()
    # groups of fewer than 4 used nodes give no columns (fe/solver.py:186-187)
    assert O.rigid_modes(g.node_xyz, nmap, K.shape[0], [np.arange(3)]).shape[1] == 0


def test_rigid_modes_per_bone_of_split_bodies():
    """Two bones meeting at a plane with no joint cells: split connectivity gives each bone its own
    nodes, so each bone's 6 modes (rb/splitfe.py:105-120 on orig coordinates) are zero-energy
    modes of the assembled split system -- 12 in all -- while the modes of the bones taken as one
    body are not."""
    g = O.Grid(np.arange(7.0), np.arange(4.0), np.arange(4.0))
    active = np.ones(g.ncell, bool)
    part = np.where(g.cI < 3, 1, 2)
    share = np.zeros((g.ncell, 3), np.float32)
    cn, orig, nsplit = O.split_connectivity_fast(g, active, part, share)
    assert nsplit == 16                            # the 4 x 4 nodes of the plane x = 3
    lam, mu = O.lame(np.full(g.ncell, 700.0), np.full(g.ncell, 0.3))
    K, nmap, used, cells, _ = O.assemble(g, lam, mu, active, KLf=g.KL_sri, cn=cn)
    xyz = g.node_xyz[orig]
    groups = [np.unique(cn[part == b]) for b in (1, 2)]
    Z = O.rigid_modes(xyz, nmap, K.shape[0], groups).toarray()
    assert Z.shape[1] == 12 and np.linalg.matrix_rank(Z) == 12
    assert np.abs(K @ Z).max() <= 1e-11 * abs(K).max()
    Zb = O.rigid_modes(xyz, nmap, K.shape[0], [np.arange(len(orig))]).toarray()
    assert np.abs(K @ Zb).max() <= 1e-11 * abs(K).max()      # one body: rigid too (both move together)
    # a relative rotation of the two bones is free (no joint), a relative stretch inside one is not
    assert np.abs(K @ (Z[:, 5] - Z[:, 11])).max() <= 1e-11 * abs(K).max()


def test_element_forces_scatter_to_the_assembled_product():
    """rb/splitfe.py:178-193: the per-cell forces K_e u_e, summed into their nodes, are K u of the
    assembled matrix (assemble_cn, pinned by the brute-force element loop) -- a transposed K_e, a
    wrong corner order or a dropped lambda term breaks the sum; SRI and full integration."""
    from comech_amd import synth
    c = synth.hindfoot_raster(dims=(12, 8, 8), h=3.0)
    og = O.Grid(c["grid"].xe, c["grid"].ye, c["grid"].ze)
    rng = np.random.default_rng(4)
    for sri in (True, False):
        K, nmap, used, cells, _ = O.assemble(og, c["lam"], c["mu"], c["active"], KLf=og.KL_sri if sri else og.KL,
                                             cn=c["cn"])
        U = np.full((len(c["orig"]), 3), np.nan)
        U[used] = rng.standard_normal((len(used), 3))
        F = O.element_forces_cn(og, c["lam"], c["mu"], U, cells, c["cn"], sri=sri)
        node_f = np.zeros((len(c["orig"]), 3))
        np.add.at(node_f, c["cn"][cells].ravel(), F.reshape(-1, 3))
        Ku = (K @ U[used].ravel()).reshape(-1, 3)
        assert np.abs(node_f[used] - Ku).max() <= 1e-12 * np.abs(Ku).max()
        # an inactive cell contributes nothing; the NaN of unused nodes counts as zero
        assert not np.isnan(F).any()


def test_penalties_are_rank_one_with_unit_weight_sum():
    """fe/solver.py:222-232: each penalty adds k w w^T with w the given weights on component comp of
    the nodes, normalised to sum 1 (duplicate nodes add their weights)."""
    g, lam, mu = _free_body(1)
    active = np.ones(g.ncell, bool)
    nodes = np.array([3, 7, 7, 11, 20])
    weights = np.array([1.0, 2.0, 0.5, 1.5, 3.0])
    k = 250.0
    _, i0 = O.solve(g, lam, mu, active, [], [], method="none")
    _, i1 = O.solve(g, lam, mu, active, [], [], method="none", penalties=[(nodes, weights, 2, k)])
    D = (i1["K"] - i0["K"]).toarray()
    w = np.zeros(D.shape[0])
    for n_, wt in zip(nodes, weights):
        w[3 * i0["nmap"][n_] + 2] += wt
    w /= weights.sum()
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs(D - k * np.outer(w, w)).max() <= 1e-13 * k
    # the penalty pulls the weighted mean of u_z: a rigid z translation t costs k t^2
    t = np.zeros(D.shape[0]); t[2::3] = 1.0
    assert abs(t @ D @ t - k) <= 1e-12 * k


def test_region_labels_rules_and_precedence():
    """rb/rbfe.py:190-201 on hand-built cells: the dominant class, then coalition/fusion, fibrous,
    interposition, then cartilage bodies override, in that order."""
    C = O.LATTICE_CLASSES
    F = np.zeros((8, 10))
    F[0, C.index("soft")] = 1.0                              # soft
    F[1, C.index("fascia")] = 0.6; F[1, C.index("soft")] = 0.4
    F[2, C.index("cortical")] = 0.5; F[2, C.index("soft")] = 0.5   # tie: the first maximum (cortical)
    F[3, C.index("cancellous")] = 1.0
    F[4, C.index("plate")] = 1.0                             # plate dominates: 'other'
    F[5, C.index("cortical")] = 1.0
    F[6, C.index("cortical")] = 1.0
    F[7, C.index("fibrous")] = 1.0
    L = np.zeros((8, 4), np.int64)
    L[5, 0] = O.FUSE                                         # bone with one fusion sample -> coalition
    L[6, 1] = O.ENGINE_CART["post"]; L[6, 2] = O.COAL        # cartilage wins over coalition
    L[7, :] = O.FIB; L[7, 3] = O.INTERP                      # interposition wins over fibrous
    lab, names = O.region_labels(F, L)
    assert names[-7:] == ["coalition", "fibrous", "fat", "soft", "fascia", "bone", "other"]
    assert [names[i] for i in lab] == ["soft", "fascia", "bone", "bone", "other", "coalition", "post", "fat"]


def test_bone_forces_uniform_stress_closed_form():
    """Two cells along x, the first the bone: under a uniform strain the second cell's element
    forces on the shared face's nodes sum to -sigma e_x (hy hz), so the force on the bone from that
    region is sigma e_x hy hz (full and one-point lambda integration alike)."""
    g = O.Grid(np.r_[0.0, 1.5, 2.5], np.r_[0.0, 0.8], np.r_[0.0, 1.2])
    lam, mu = np.array([3.0, 2.0]), np.array([1.0, 0.7])
    G = np.array([[1e-3, 2e-4, -3e-4], [4e-4, -5e-4, 1e-4], [-2e-4, 3e-4, 6e-4]])
    U = g.node_xyz @ G.T
    eps = 0.5 * (G + G.T)
    part = np.array([1, 0])
    lab = np.array([0, 3])
    names = ["a", "b", "c", "d"]
    for sri in (True, False):
        f = O.bone_forces(g, lam, mu, U, g.cell_nodes, part, np.ones(2, bool), 1, lab, names, g.nnode, sri=sri)
        sig = lam[1] * np.trace(eps) * np.eye(3) + 2 * mu[1] * eps
        assert list(f) == ["d"]
        assert np.allclose(f["d"], sig[:, 0] * 0.8 * 1.2, rtol=1e-12, atol=1e-16)


def test_ground_zones_hand_built():
    R = np.zeros((6, 3))
    R[:, 2] = [1.0, 2.0, 4.0, 8.0, 16.0, 32.0]
    x = np.array([-50.0, -31.0, -30.0, 0.0, 25.0, 26.0])
    z = O.ground_zones(R, x, scale=0.75, c0x=10.0)        # xh = 10 + 0.75 (-50) = -27.5, xf = 10 + 0.75 * 10 = 17.5
    assert z == {"heel": 7.0, "mid": 8.0, "fore": 48.0}


def test_rigid_fit_recovers_rigid_motion():
    rng = np.random.default_rng(4)
    X = rng.uniform(-20, 20, (50, 3)) + [3.0, -1.0, 7.0]
    t, w = np.array([0.1, -0.2, 0.05]), np.array([1e-3, -2e-3, 5e-4])
    c = X.mean(0)
    U = t + np.cross(w, X - c)
    tt, ww, cc = O.rigid_fit(X, U)
    assert np.allclose(tt, t, atol=1e-14) and np.allclose(ww, w, atol=1e-15) and np.allclose(cc, c)


def test_effects_hand_built():
    """A 6x4x3 grid: von Mises known per cell, cartilage/coalition bodies in chosen cells; the talus
    moves rigidly relative to the calcaneus (translation, rotation about x); checks every descriptor
    against its definition."""
    g = O.Grid(np.arange(7.0), np.arange(5.0) * 0.5, np.arange(4.0))
    nc = g.ncell
    rng = np.random.default_rng(8)
    VM = rng.uniform(0, 10, nc).astype(np.float32)
    active = np.ones(nc, bool)
    active[5] = False
    L = np.zeros((nc, 4), np.int64)
    ant = np.arange(10, 20)
    L[ant, 0] = O.ENGINE_CART["ant"]                          # fraction 0.25 -> in p99 and mean
    L[ant[:3], 1] = O.ENGINE_CART["ant"]                      # fraction 0.5
    L[30, 0] = O.ENGINE_CART["mid"]                           # a single mid cell: p99 = 0 (< 3 cells)
    coal = np.arange(40, 48)
    L[coal, :2] = O.COAL
    part = np.zeros(nc, np.int64)
    bnames = {1: "calcaneus", 2: "talus", 3: "navicular"}
    xyz = g.node_xyz
    calc, tal, nav = np.arange(0, 30), np.arange(30, 60), np.arange(60, 70)
    base_t, base_w = np.array([0.01, 0.02, -0.05]), np.array([1e-3, 2e-4, -3e-4])
    c_c = xyz[calc].mean(0)
    U = base_t + np.cross(base_w, xyz - c_c)                  # everything moves with the calcaneus
    rel_t, rel_w = np.array([0.0, 0.003, -0.02]), np.array([0.004, 0.0, 0.0])
    c_t = xyz[tal].mean(0)
    U[tal] += rel_t + np.cross(rel_w, xyz[tal] - c_t)
    U[nav, 2] -= 0.01
    bone_nodes = {"calcaneus": calc, "talus": tal, "navicular": nav}
    Ft = {"ant": [0, 0, 30.0], "mid": [0, 0, -5.0], "post": [1, 1, 10.0], "soft": [0, 0, 99.0]}
    gz = {"heel": 150.0, "mid": 50.0, "fore": 100.0}
    d, Urel = O.effects(VM, L, active, part, bnames, g.cell_vol, U, xyz, bone_nodes, Ft, gz, 300.0)
    sel = ant[active[ant]]
    assert np.isclose(d["ant_p99"], np.percentile(VM[sel], 99))
    w = np.where(np.isin(sel, ant[:3]), 0.5, 0.25) * g.cell_vol[sel]
    assert np.isclose(d["ant_mean"], (VM[sel] * w).sum() / w.sum())
    assert d["mid_p99"] == 0.0 and np.isclose(d["mid_mean"], VM[30])
    assert d["post_p99"] == 0.0 and d["post_mean"] == 0.0
    assert np.isclose(d["share_ant"], 30 / 40) and d["share_mid"] == 0.0 and np.isclose(d["share_post"], 10 / 40)
    assert np.isclose(d["talus_dz"], rel_t[2]) and np.isclose(d["talus_dy"], rel_t[1])
    assert np.isclose(d["talus_evert"], np.rad2deg(rel_w[0]))
    assert np.isclose(d["nav_drop"], 0.01)
    assert np.isclose(d["heel_share"], 0.5)
    cm = coal[active[coal]]
    assert np.isclose(d["coal_p99"], np.percentile(VM[cm], 99)) and d["coal_max"] == VM[cm].max()
    assert "tc_p99" not in d
    assert np.abs(Urel[calc]).max() < 1e-15
