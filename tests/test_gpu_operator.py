"""GPU parity of the matrix-free operator, the coarse Galerkin levels and the V-cycle against the
FP64 oracle's assembled matrices (same seeded inputs).  Tolerances: FP64 relative max error
1e-12 (FP rounding of a different summation order), FP32 1e-5."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from comech_amd import synth
from comech_amd.grid import Grid
from comech_amd.system import VoxelSystem
from oracle import fe_oracle as O
from tests import cases as C

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _sys1(n=10, **kw):
    c = C.cube(n)
    extra, f = synth.solve_inputs_cube(c)
    return c, VoxelSystem(c["grid"], sri=False, active=c["active"], labels=c["labels"], lam_tab=c["lam_tab"],
                          mu_tab=c["mu_tab"], extra=extra, device=DEV, **kw)


def _sys2(dims=(40, 24, 24), h=2.0, as_terms=True, **kw):
    c = C.hindfoot(dims, h)
    extra, terms, f = synth.solve_inputs_hindfoot(c, as_terms=as_terms)
    irr_c, irr_n = __import__("comech_amd.engine", fromlist=["x"]).split_cells(c["grid"], c["active"], c["cn"])
    s = VoxelSystem(c["grid"], sri=True, active=c["active"], labels=c["labels"], lam_tab=c["lam_tab"],
                    mu_tab=c["mu_tab"], copy_orig=c["orig"][c["grid"].nnode:], irr_cells=irr_c, irr_nodes=irr_n,
                    extra=extra, rank1=terms, device=DEV, **kw)
    return c, s


def _check_apply(s, info, nn, tol64=1e-12, tol32=1e-5):
    used = np.zeros(nn, bool)
    used[info["used"]] = True
    u = C.random_full(nn, used, seed=1)
    K = info["K"]
    ref = K @ u.reshape(-1, 3)[info["used"]].ravel()
    for dt, tol in ((torch.float64, tol64), (torch.float32, tol32)):
        y = s.apply(torch.tensor(u, dtype=dt, device=DEV)).double().cpu().numpy().reshape(-1, 3)
        got = y[info["used"]].ravel()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < tol, (dt, err)
        # rows of unused nodes are empty
        assert not np.any(y[~used])


def test_apply_cube_inclusions_full_integration():
    c, s = _sys1(10)
    _check_apply(s, C.oracle_system1(10), c["grid"].nnode)


@pytest.mark.parametrize("as_terms", [True, False])
def test_apply_hindfoot_raster_split_sri_springs(as_terms):
    """Springs as rank-one terms or as generic COO (diagonal -> rank-one, off-diagonal -> CSR)."""
    c, s = _sys2(as_terms=as_terms)
    assert s.ncopy > 0 and len(s.irr_cells) > 0
    _check_apply(s, C.oracle_system2(), len(c["orig"]))


def test_apply_rectilinear_value_mode_odd_sizes():
    """Stretched grid (fe/test_patch.py:4 edges), per-cell random materials (value mode: more than
    255 distinct pairs), inactive cells, both integrations."""
    rng = np.random.default_rng(7)
    xe = np.r_[0, 1, 3, 4, 7, 10, 10.5]
    ye = np.r_[0, 2, 5, 10, 11]
    ze = np.r_[0, 1, 2, 4, 6, 10, 11, 13]
    g, og = Grid(xe, ye, ze), O.Grid(xe, ye, ze)
    lam = rng.uniform(1, 100, g.ncell)
    mu = rng.uniform(1, 100, g.ncell)
    active = rng.uniform(size=g.ncell) > 0.2
    for sri in (False, True):
        s = VoxelSystem(g, sri=sri, active=active, lam=lam, mu=mu, device=DEV, min_dofs=10, materials="values")
        assert not s.label_mode
        K, nmap, used, cells, _ = O.assemble(og, lam, mu, active, KLf=og.KL_sri if sri else og.KL)
        _check_apply(s, dict(K=K, used=used), g.nnode)


def _oracle_hierarchy(K, used, grid, levels):
    mg = O.GMG(K, used, grid.xe, grid.ye, grid.ze, levels=levels)
    return mg


def _patch_case():
    """fe/test_patch.py grid: stretched edges, odd node counts (a last coarse cell one fine cell
    wide), springs spanning 1e-6 .. 1e9 per area; 6 levels (axes of one cell keep their size)."""
    xe, ye, ze = np.r_[0, 1, 3, 4, 7, 10], np.r_[0, 2, 5, 10], np.r_[0, 1, 2, 4, 6, 10]
    g, og = Grid(xe, ye, ze), O.Grid(xe, ye, ze)
    lam, mu = O.lame(np.full(g.ncell, 1000.0), np.full(g.ncell, 0.3))
    bot, top = np.flatnonzero(og.cK == 0), np.flatnonzero(og.cK == og.nz - 1)
    nb, ab = O.face_node_areas(og, bot, "z-")
    nt, at = O.face_node_areas(og, top, "z+")
    springs = [(nb, ab, np.array([1e-6, 1e-6, 1e9]))]
    _, info = O.solve(og, lam, mu, np.ones(g.ncell, bool), springs, [(nt, at, np.array([0, 0, -1.0]))],
                      method="none")
    kd = np.zeros((g.nnode, 3))
    np.add.at(kd, nb, ab[:, None] * springs[0][2][None, :])
    rows = np.arange(3 * g.nnode)[kd.ravel() != 0]
    s = VoxelSystem(g, sri=False, active=np.ones(g.ncell, bool), lam=lam, mu=mu, extra=(rows, rows, kd.ravel()[rows]),
                    device=DEV, min_dofs=10)
    return dict(grid=g), s, info, info["used"]


@pytest.mark.parametrize("which", ["cube", "hindfoot", "hindfoot_coo", "patch", "hindfoot_semi"])
def test_coarse_levels_equal_galerkin(which):
    """Level-l GPU operator (moment cells + Galerkin springs) == P^T A P of the oracle hierarchy
    (fe/gmg.py:91-99) on the kept coarse DOFs, every level.  hindfoot_semi: automatic level rule with the
    last step coarsening x and y only (the oracle's prolongation with the same axis mask)."""
    nlev = 3
    if which == "hindfoot_semi":
        c, s = _sys2(last_axes=3)
        info = C.oracle_system2()
        used_nodes = c["orig"][info["used"]]
        g = c["grid"]
        mg = O.GMG(info["K"], used_nodes, g.xe, g.ye, g.ze, last_axes=3)
        nlev = len(mg.levels)
        assert nlev >= 3 and len(s.levels()) == nlev
        top = s.levels()[-1]
        assert top["nz"] == s.levels()[-2]["nz"] and top["nx"] < s.levels()[-2]["nx"]
    elif which == "cube":
        c, s = _sys1(10, levels=3)
        info = C.oracle_system1(10)
        used_nodes = info["used"]
    elif which == "patch":
        c, s, info, used_nodes = _patch_case()
        nlev = 6
    else:
        c, s = _sys2(levels=3, as_terms=(which == "hindfoot"))
        info = C.oracle_system2()
        used_nodes = c["orig"][info["used"]]
    g = c["grid"]
    if which != "hindfoot_semi":
        mg = _oracle_hierarchy(info["K"], used_nodes, g, nlev)
    lv = s.levels()
    assert len(lv) == nlev
    if which.startswith("hindfoot") or which == "cube":
        # level 1 runs both of its paths: cells of one material in the structured march, cells of
        # several materials in the heterogeneous-cell pass
        st = s.level_stats(1)
        assert 0 < st["het_cells"] < st["cells"], st
    for l in range(1, nlev):
        Ac = mg.levels[l].A
        keep = mg.keeps[l - 1]
        nn = lv[l]["nodes"]
        rng = np.random.default_rng(l)
        u = np.zeros(3 * nn)
        u[keep] = rng.standard_normal(len(keep))
        y = s.apply(torch.tensor(u, dtype=torch.float64, device=DEV), level=l).cpu().numpy()
        ref = Ac @ u[keep]
        err = np.abs(y[keep] - ref).max() / np.abs(ref).max()
        assert err < 1e-11, (l, err)
        assert lv[l]["ndof"] == len(keep)
        # a rigid x translation: elasticity gives nothing, only the (possibly weak) springs respond,
        # so a dropped or mis-coarsened spring term shows up here however small it is
        u = np.zeros(3 * nn)
        u[keep[keep % 3 == 0]] = 1.0
        y = s.apply(torch.tensor(u, dtype=torch.float64, device=DEV), level=l).cpu().numpy()
        ref = Ac @ u[keep]
        err = np.abs(y[keep] - ref).max() / np.abs(ref).max()
        assert err < 1e-6, (l, "rigid", err)


@pytest.mark.parametrize("last_axes,expect", [(7, [107811, 14739, 2187]), (3, [107811, 14739, 4131])])
def test_level_rule_matches_reference(last_axes, expect):
    """<= 6000 DOFs or 6 levels (fe/gmg.py:80,87-90): 32^3 gives 3 levels (107,811 -> 14,739 ->
    2,187 DOFs) under the reference rule (every axis coarsened).  The default semi-coarsening of
    the last step keeps z: 8 x 8 x 16 cells, 9 * 9 * 17 * 3 = 4,131 DOFs (still <= 6000)."""
    c = synth.cube_inclusions(n=32)
    extra, _ = synth.solve_inputs_cube(c)
    s = VoxelSystem(c["grid"], sri=False, active=c["active"], labels=c["labels"], lam_tab=c["lam_tab"],
                    mu_tab=c["mu_tab"], extra=extra, device=DEV, last_axes=last_axes)
    assert [l["ndof"] for l in s.levels()] == expect
    if last_axes == 3:
        assert (s.levels()[-1]["nx"], s.levels()[-1]["ny"], s.levels()[-1]["nz"]) == (8, 8, 16)


@pytest.mark.parametrize("smoother", ["chebyshev", "jacobi"])
def test_vcycle_is_spd_and_contracts(smoother):
    """The V-cycle is a symmetric positive definite preconditioner (required by PCG) and
    I - B A contracts on the fine level."""
    c, s = _sys2(smoother=smoother)
    info = C.oracle_system2()
    nn = len(c["orig"])
    used = np.zeros(nn, bool)
    used[info["used"]] = True
    r1, r2 = C.random_full(nn, used, 3), C.random_full(nn, used, 4)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=DEV)
    z1, z2 = s.vcycle(t(r1)).cpu().numpy(), s.vcycle(t(r2)).cpu().numpy()
    assert abs(r2 @ z1 - r1 @ z2) <= 1e-9 * abs(r1 @ z2)
    assert r1 @ z1 > 0
    e = C.random_full(nn, used, 5)
    Ae = s.apply(t(e))
    e1 = e - s.vcycle(Ae).cpu().numpy()
    K = info["K"]
    en = lambda v: np.sqrt(v.reshape(-1, 3)[info["used"]].ravel() @ (K @ v.reshape(-1, 3)[info["used"]].ravel()))
    assert en(e1) < 0.7 * en(e)


def test_single_level_vcycle_is_dense_solve():
    """One level: the V-cycle is the dense coarsest solve (blocked Gauss-Jordan inverse of the
    3195-DOF split-node system, 50 pivot blocks, the last one partial) and equals K^-1 r of the
    oracle's assembled matrix (cond(K) ~ 1e6, so 1e-8 leaves room for FP64 rounding)."""
    import scipy.sparse.linalg as sla
    c, s = _sys2(dims=(12, 8, 8), h=3.0, levels=1)
    info = C.oracle_system2((12, 8, 8), 3.0)
    nn = len(c["orig"])
    used = np.zeros(nn, bool)
    used[info["used"]] = True
    r = C.random_full(nn, used, 7)
    z = s.vcycle(torch.tensor(r, dtype=torch.float64, device=DEV)).cpu().numpy().reshape(-1, 3)
    ref = sla.spsolve(info["K"].tocsc(), r.reshape(-1, 3)[info["used"]].ravel())
    got = z[info["used"]].ravel()
    assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()
    assert not np.any(z[~used])


@pytest.mark.parametrize("dt,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_vcycle_graph_replays_match_eager(dt, tol):
    """The V-cycle replayed from its captured HIP graph equals the eagerly launched V-cycle on
    every replay, with other work (operator applications, torch kernels, host synchronisation)
    between the replays.  Tolerance: rounding level (the sparse pass accumulates with atomics)."""
    c, s = _sys2(dims=(24, 16, 16), h=3.0)
    nn = len(c["orig"])
    rng = np.random.default_rng(3)
    r = torch.tensor(rng.standard_normal(3 * nn), dtype=dt, device=DEV)
    s.set_graphs(False)
    ref = s.vcycle(r).clone()
    s.set_graphs(True)
    scale = ref.abs().max().item()
    z = torch.empty_like(r)                       # same buffers on every call: one capture, then replays
    for rep in range(4):
        z.fill_(float("nan"))
        s.vcycle(r, z)
        err = (z - ref).abs().max().item() / scale
        assert err <= tol, (rep, err)
        # eager work between replays
        y = s.apply(r)
        _ = (y * 2.0 + r).sum().item()
        torch.cuda.synchronize()


@pytest.mark.parametrize("kz", [0, 1, 4])
def test_block_skip_equals_full_march(kz):
    """Structured-march blocks whose cells carry no stiffness run only the epilogue; the result
    equals the full march on every level, in both precisions (operator and V-cycle), and FP32
    equals FP64 to FP32 rounding (covers the level-1 child-weight and coarser moment cells)."""
    c, s = _sys2(levels=4)
    if kz:
        s.set_kz(kz)
    lv = s.levels()
    rng = np.random.default_rng(5)
    for l in range(len(lv) - 1):
        u = rng.standard_normal(3 * lv[l]["nodes"])
        outs = {}
        for skip in (True, False):
            s.set_block_skip(skip)
            for dt in (torch.float64, torch.float32):
                outs[skip, dt] = s.apply(torch.tensor(u, dtype=dt, device=DEV), level=l).double().cpu().numpy()
        ref = outs[False, torch.float64]
        scale = np.abs(ref).max()
        assert np.abs(outs[True, torch.float64] - ref).max() <= 1e-13 * scale, l
        assert np.abs(outs[True, torch.float32] - outs[False, torch.float32]).max() <= 1e-6 * scale, l
        assert np.abs(outs[True, torch.float32] - ref).max() <= 1e-5 * scale, l
    nn = len(c["orig"])
    r = rng.standard_normal(3 * nn)
    zs = {}
    for skip in (True, False):
        s.set_block_skip(skip)
        for dt in (torch.float64, torch.float32):
            zs[skip, dt] = s.vcycle(torch.tensor(r, dtype=dt, device=DEV)).double().cpu().numpy()
    scale = np.abs(zs[False, torch.float64]).max()
    assert np.abs(zs[True, torch.float64] - zs[False, torch.float64]).max() <= 1e-12 * scale
    assert np.abs(zs[True, torch.float32] - zs[False, torch.float32]).max() <= 1e-5 * scale
    s.close()
