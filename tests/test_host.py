"""CPU tests of the product's host logic and of the kernel math (CPU build of element.hpp), plus
the C-ABI library: it loads and exports every symbol include/comech_fe.h declares (no compute calls
without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from comech_amd import _lib, synth
from comech_amd import engine as E
from comech_amd.grid import Grid
from comech_amd.rb import gpusolve as GS
from comech_amd.rb import splitfe as SF
from comech_amd.system import compress_materials
from oracle import fe_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for header, table in (("comech_fe.h", _lib.SIGNATURES), ("comech_learned.h", _lib.SIGNATURES_NN)):
        hdr = open(os.path.join(ROOT, "include", header)).read()
        names = set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+(cm_\w+)\s*\(", hdr, re.M))
        assert len(names) >= 15, header
        for n in sorted(names):
            assert hasattr(lib, n), n
        assert names == set(table), names ^ set(table)
    assert lib.cm_version() == 1


def test_library_error_path_without_gpu():
    """Argument errors come back as status -1 with a message (no device touched)."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    e = np.zeros(1)
    rc = lib.cm_sys_create(ctypes.byref(h), None, 0, 1, 1, _lib.ptr(e, ctypes.c_double), _lib.ptr(e, ctypes.c_double),
                           _lib.ptr(e, ctypes.c_double), 1)
    assert rc == -1 and b"at least one cell" in lib.cm_last_error()


def _host():
    lib = ctypes.CDLL(os.path.join(ROOT, "comech_amd", "libcomech_host.so"))
    return lib


P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("dims", [(1, 1, 1), (0.5, 0.7, 1.3), (2.0, 3.0, 0.25)])
@pytest.mark.parametrize("sri", [0, 1])
def test_kernel_element_operator_matches_oracle(dims, sri):
    """The Walsh-Hadamard-basis operator of the kernels equals lam*KL + mu*KM (fe/solver.py:77-90,
    rb/splitfe.py:58-69), FP64 to rounding and FP32 to 1e-6."""
    lib = _host()
    KL, KM, _ = O.element_matrices(dims)
    if sri:
        KL = O.kl_sri(dims)
    K = np.zeros((24, 24))
    lib.cm_host_elem_const(P(np.asarray(dims, float)), ctypes.c_double(1.7), ctypes.c_double(2.3), sri, P(K))
    ref = 1.7 * KL + 2.3 * KM
    assert np.abs(K - ref).max() <= 1e-14 * np.abs(ref).max()
    u = np.random.default_rng(0).standard_normal(24)
    f = np.zeros(24)
    lib.cm_host_elem_const_f32(P(np.asarray(dims, float)), ctypes.c_double(1.7), ctypes.c_double(2.3), sri, P(u), P(f))
    assert np.abs(f - ref @ u).max() <= 1e-5 * np.abs(ref @ u).max()


@pytest.mark.parametrize("sri", [0, 1])
@pytest.mark.parametrize("iso", [0, 1])
def test_march_column_equals_assembled_operator(sri, iso):
    """The structured kernel's z-march (face transforms, z stage, transform-space cell operator,
    force carry, corner back-transform), run on one cell column on the CPU, equals K u of the
    oracle's assembly (fe/solver.py:137-162, SRI rb/splitfe.py:58-69).  iso=1 is the cubic-cell
    form with the gradient factors folded into the material weights."""
    lib = _host()
    rng = np.random.default_rng(10 * sri + iso)
    nz = 5
    if iso:
        hxy = np.array([0.5, 0.5])
        hz = np.full(nz, 0.5)
    else:
        hxy = np.array([0.7, 1.3])
        hz = np.round(rng.uniform(0.3, 1.5, nz), 2)        # the oracle groups cells by rounded dims
    g = O.Grid(np.r_[0.0, hxy[0]], np.r_[0.0, hxy[1]], np.r_[0.0, np.cumsum(hz)])
    hz = g.dz
    lam, mu = rng.uniform(0.1, 5, nz), rng.uniform(0.1, 5, nz)
    lam[2] = mu[2] = 0.0                                   # an empty cell inside the column
    K, nmap, used, _, _ = O.assemble(g, lam, mu, np.ones(nz, bool), KLf=g.KL_sri if sri else g.KL)
    u = rng.standard_normal(3 * g.nnode)
    y = np.zeros(3 * g.nnode)
    lib.cm_host_march_column(nz, P(hxy), P(hz), P(lam), P(mu), sri, iso, P(u), P(y))
    ref = np.zeros(3 * g.nnode)
    ud = (3 * used[:, None] + np.arange(3)).ravel()
    ref[ud] = K @ u[ud]
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()


def test_face_form_equals_transform_composition():
    """The level-0 kernel's face form of one z step (cell_faces_iso_sri: z butterflies folded into
    the element terms, normal and shear weights of a mode combined) equals the composition it
    replaces (faces_to_cell, the transform-space operator, cell_to_faces, carry add) on random
    faces, and a zero material gives back the carry."""
    lib = _host()
    rng = np.random.default_rng(7)
    for trial in range(200):
        lo, hi, car = (rng.standard_normal(12) * 10.0 ** rng.uniform(-3, 3) for _ in range(3))
        wl, wm = 10.0 ** rng.uniform(-3, 3, 2)
        out = {}
        for fused in (0, 1):
            flo, fhi = np.zeros(12), np.zeros(12)
            lib.cm_host_cell_faces(fused, P(lo), P(hi), P(car), ctypes.c_double(wl), ctypes.c_double(wm), P(flo), P(fhi))
            out[fused] = np.r_[flo, fhi]
        scale = np.abs(out[0]).max()
        assert np.abs(out[1] - out[0]).max() <= 1e-14 * scale, trial
    flo, fhi = np.zeros(12), np.zeros(12)
    lib.cm_host_cell_faces(1, P(lo), P(hi), P(car), ctypes.c_double(0.0), ctypes.c_double(0.0), P(flo), P(fhi))
    assert np.array_equal(flo, car) and not fhi.any()


@pytest.mark.parametrize("sri", [0, 1])
@pytest.mark.parametrize("split", [(2, 2, 2), (2, 1, 2), (1, 2, 1)])
def test_homogeneous_level1_cell_operator(sri, split):
    """A level-1 cell whose level-0 children (equal halves along each split axis, one child along
    the others) share one material: the constant-material operator with per-axis lam second
    moments q (1/4 per split axis and 0 otherwise under SRI, 1/3 with exact integration) equals the
    moment operator of the children's shifted moments, which is the exact Galerkin operator
    (test_kernel_moments_are_exact_galerkin pins that one against P^T K P)."""
    lib = _host()
    h = np.array([1.3, 0.7, 2.1])
    lam, mu = 2.7, 1.9
    momL, momM = np.zeros(27), np.zeros(27)
    V = float(np.prod(h))
    Vc = V / np.prod(split)
    for dz in range(split[2]):
        for dy in range(split[1]):
            for dx in range(split[0]):
                d = (dx, dy, dz)
                a = np.array([1.0 / n for n in split])
                b = np.array([0.0 if n == 1 else (-0.5 if d[i] == 0 else 0.5) for i, n in enumerate(split)])
                cl, cm = np.zeros(27), np.zeros(27)
                lib.cm_host_mom_const(ctypes.c_double(lam), ctypes.c_double(Vc), sri, P(cl))
                lib.cm_host_mom_const(ctypes.c_double(mu), ctypes.c_double(Vc), 0, P(cm))
                lib.cm_host_mom_shift_add(P(cl), P(a), P(b), P(momL))
                lib.cm_host_mom_shift_add(P(cm), P(a), P(b), P(momM))
    Kmom, Kq = np.zeros((24, 24)), np.zeros((24, 24))
    lib.cm_host_elem_mom(P(momL), P(momM), P(h), P(Kmom))
    q = np.array([(0.25 if n == 2 else 0.0) if sri else 1.0 / 3.0 for n in split])
    lib.cm_host_elem_q(P(h), ctypes.c_double(lam), ctypes.c_double(mu), P(q), P(Kq))
    assert np.abs(Kq - Kmom).max() <= 1e-13 * np.abs(Kmom).max()
    assert np.abs(Kq - Kq.T).max() <= 1e-13 * np.abs(Kq).max()


@pytest.mark.parametrize("sri", [0, 1])
def test_kernel_moments_are_exact_galerkin(sri):
    """Coarse cell operator from shifted child moments == P^T K P of its 2x2x2 children with random
    materials on a stretched cell (fe/gmg.py:91-97)."""
    lib = _host()
    rng = np.random.default_rng(sri)
    xe, ye, ze = np.r_[0, 0.7, 2.0], np.r_[0, 1.1, 1.5], np.r_[0, 0.4, 1.2]
    g = O.Grid(xe, ye, ze)
    lam, mu = rng.uniform(0.1, 3, 8), rng.uniform(0.1, 3, 8)
    KLf = g.KL_sri if sri else g.KL
    Kf = np.zeros((81, 81))
    mL, mM = np.zeros(27), np.zeros(27)
    H = np.array([xe[-1], ye[-1], ze[-1]])
    for c in range(8):
        d = (3 * g.cell_nodes[c][:, None] + np.arange(3)).ravel()
        Kf[np.ix_(d, d)] += lam[c] * KLf[g.dim_id[c]] + mu[c] * g.KM[g.dim_id[c]]
        lo = np.array([xe[g.cI[c]], ye[g.cJ[c]], ze[g.cK[c]]])
        hi = np.array([xe[g.cI[c] + 1], ye[g.cJ[c] + 1], ze[g.cK[c] + 1]])
        a = (hi - lo) / H
        b = (lo + hi - H) / H
        t = np.zeros(27)
        lib.cm_host_mom_const(ctypes.c_double(lam[c]), ctypes.c_double(np.prod(hi - lo)), sri, P(t))
        lib.cm_host_mom_shift_add(P(t), P(a), P(b), P(mL))
        lib.cm_host_mom_const(ctypes.c_double(mu[c]), ctypes.c_double(np.prod(hi - lo)), 0, P(t))
        lib.cm_host_mom_shift_add(P(t), P(a), P(b), P(mM))
    Px, _ = O.p1d(xe); Py, _ = O.p1d(ye); Pz, _ = O.p1d(ze)
    import scipy.sparse as sp
    Pd = np.kron(sp.kron(Pz, sp.kron(Py, Px)).toarray(), np.eye(3))
    ref = Pd.T @ Kf @ Pd
    K = np.zeros((24, 24))
    lib.cm_host_elem_mom(P(mL), P(mM), P(H), P(K))
    assert np.abs(K - ref).max() <= 1e-13 * np.abs(ref).max()


def test_kernel_cell_stress_matches_oracle():
    lib = _host()
    dims = np.array([0.5, 0.7, 1.3])
    u = np.random.default_rng(2).standard_normal(24)
    eps, sig, vm = np.zeros(6), np.zeros(6), np.zeros(1)
    lib.cm_host_cell_stress(P(u), P(dims), ctypes.c_double(3.0), ctypes.c_double(2.0), P(eps), P(sig), P(vm))
    Bc = O.element_matrices(dims)[2]
    e_ref = Bc @ u
    tr = e_ref[:3].sum()
    s_ref = np.r_[3.0 * tr + 4.0 * e_ref[:3], 2.0 * e_ref[3:]]
    assert np.allclose(eps, e_ref) and np.allclose(sig, s_ref) and np.isclose(vm[0], O.von_mises(s_ref[None])[0])


def test_product_grid_matches_oracle():
    xe, ye, ze = np.r_[0, 1, 3, 4, 7, 10], np.r_[0, 2, 5, 10], np.r_[0, 1, 2, 4, 6, 10]
    g, o = Grid(xe, ye, ze), O.Grid(xe, ye, ze)
    assert np.array_equal(g.cell_nodes, o.cell_nodes) and np.array_equal(g.dim_id, o.dim_id)
    assert np.array_equal(g.node_xyz, o.node_xyz) and np.array_equal(g.cI, o.cI)
    for a, b in ((g.KL, o.KL), (g.KM, o.KM), (g.Bc, o.Bc), (g.KL_sri, o.KL_sri), (g.KB, o.KB)):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("seed", range(6))
def test_product_split_connectivity_bit_exact(seed):
    """Node numbering, split copies and orig: product == oracle loop form, bit for bit."""
    rng = np.random.default_rng(100 + seed)
    shape = (7, 5, 6)
    g, og = Grid(*(np.arange(n + 1.0) for n in shape)), O.Grid(*(np.arange(n + 1.0) for n in shape))
    part = rng.integers(0, 4, g.ncell)
    active = rng.uniform(size=g.ncell) > 0.1
    share = rng.uniform(size=(g.ncell, 4)).astype(np.float32)
    share[rng.uniform(size=g.ncell) < 0.3] = 0
    share[:, 0] = 0
    cn, orig, ns = SF.split_connectivity_fast(g, active, part, share)
    cno, origo, nso = O.split_connectivity(og, active, part, share)
    assert ns == nso and np.array_equal(orig, origo) and np.array_equal(cn[active], cno[active])
    assert cn.dtype == np.int64 and orig.dtype == np.int64


def test_used_nodes_and_irregular_cells():
    c = synth.hindfoot_raster(dims=(20, 12, 12), h=4.0)
    g = c["grid"]
    irr_c, irr_n = E.split_cells(g, c["active"], c["cn"])
    used = E.used_nodes(g, c["active"], len(c["orig"]), irr_c, irr_n)
    ref = np.zeros(len(c["orig"]), bool)
    ref[np.unique(c["cn"][c["active"]])] = True
    assert np.array_equal(used, ref)
    assert np.array_equal(irr_c, np.flatnonzero(c["active"] & (c["cn"] != g.cell_nodes).any(1)))


def test_compress_materials_exact():
    rng = np.random.default_rng(3)
    lam = rng.choice([1.0, 2.0, 3.0], 100)
    mu = rng.choice([0.5, 0.25], 100)
    act = rng.uniform(size=100) > 0.3
    lt, mt, lab = compress_materials(lam, mu, act)
    assert np.array_equal(lt[lab[act]], lam[act]) and np.array_equal(mt[lab[act]], mu[act])
    assert (lab[~act] == 255).all()
    assert compress_materials(rng.uniform(size=1000), rng.uniform(size=1000), np.ones(1000, bool)) is None


def test_synth_deterministic_and_sized():
    a, b = synth.hindfoot_raster(dims=(24, 16, 16), h=3.0), synth.hindfoot_raster(dims=(24, 16, 16), h=3.0)
    for k in ("labels", "cn", "orig", "active"):
        assert np.array_equal(a[k], b[k])
    assert a["n_lig"] > 0 and a["n_copies"] > 0
    r, c_, v = a["extra"]
    assert len(r) == len(c_) == len(v)
    c1 = synth.cube_inclusions(n=32)
    assert c1["grid"].ncell == 32 ** 3 and set(np.unique(c1["labels"])) == {0, 1}
    sizes = synth.batch_sizes()
    assert len(sizes) == 64 and all(96 <= s[0] <= 192 and 64 <= s[1] <= 112 and 64 <= s[2] <= 112 for s in sizes)


def test_rigid_modes_match_oracle():
    g = Grid(np.arange(4.0), np.arange(3.0), np.arange(3.0))
    og = O.Grid(g.xe, g.ye, g.ze)
    used = np.ones(g.nnode, bool)
    used[5] = False
    nmap = -np.ones(g.nnode, np.int64)
    nmap[used] = np.arange(used.sum())
    groups = [np.arange(0, 20), np.arange(15, 30), np.arange(3)]
    from comech_amd.fe.solver import rigid_modes
    Zp = rigid_modes(g, nmap, 3 * used.sum(), groups).toarray()
    Zo = O.rigid_modes(og.node_xyz, nmap, 3 * used.sum(), groups).toarray()
    assert np.allclose(Zp, Zo)


def test_gpusolve_failure_hook_raises_without_cpu_fallback(monkeypatch):
    """COMECH_TEST_GPU_FAIL fires before any device call (rb/gpusolve.py:207-208); after the
    retries the product raises instead of solving on the CPU."""
    import scipy.sparse as sp
    monkeypatch.setenv("COMECH_TEST_GPU_FAIL", "1")
    monkeypatch.setenv("COMECH_RETRY_SLEEP", "0")
    with pytest.raises(RuntimeError, match="simulated GPU failure"):
        GS.pcg_gmg_gpu(sp.identity(3, format="csr"), np.ones(3), np.array([0]), (np.arange(2.0),) * 3,
                       device="cuda:0")


def test_hindfoot_family_equals_separate_generation():
    """synth.hindfoot_family computes the anatomy once and the split/spring topology once for the
    actions that keep the baseline's cells; every action's arrays equal a separate generation
    (whose topology is computed from its own cells)."""
    from comech_amd import synth
    fam = synth.hindfoot_family((24, 16, 16), 3.0, 2100, angle_seed=2600)
    keys = ("lam", "mu", "active", "labels", "lam_tab", "mu_tab", "part", "cn", "orig", "used", "terms", "extra",
            "fixed", "loads")

    def same(a, b):
        if isinstance(a, np.ndarray):
            return a.dtype == b.dtype and np.array_equal(a, b)
        if isinstance(a, (tuple, list)):
            return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
        return a == b
    for a in synth.ACTIONS:
        one = synth.hindfoot_raster(dims=(24, 16, 16), h=3.0, seed=2100, action=a, angle_seed=2600)
        for k in keys:
            assert same(fam[a][k], one[k]), (a, k)
    assert fam["fusion"]["terms"][3].max() == 1e6 and fam["baseline"]["terms"][3].max() < 1e6
    assert not np.array_equal(fam["osteotomy"]["active"], fam["baseline"]["active"])


def test_fragment_groups_equal_a_loop_over_cells():
    """engine.fragment_groups (the pieces of cut bones, vectorised inside each bone's bounding box)
    equals the direct form: label every bone's active cells over the whole grid, then take each
    piece's cells one by one with the explicit nodes of the irregular cells; the osteotomy splits
    one bone into two pieces, the baseline none."""
    import scipy.ndimage as ndi
    from comech_amd import engine as E
    from comech_amd.grid import Grid, corners_of
    for action, npieces in (("osteotomy", 2), ("baseline", 0)):
        d = synth.batch_foot(0, scale=4, actions=[action])[0]
        g = Grid(*d["edges"])
        act = np.asarray(d["labels"]) != 255
        got = E.fragment_groups(g, d["part"], act, d["irr_cells"], d["irr_nodes"])
        part = np.asarray(d["part"]).reshape(g.nx, g.ny, g.nz)
        irr = {int(c): i for i, c in enumerate(d["irr_cells"])}
        ref = []
        for p in np.unique(part[(part > 0) & act.reshape(part.shape)]):
            comp, n = ndi.label((part == p) & act.reshape(part.shape))
            if n < 2:
                continue
            for q in range(1, n + 1):
                cells = np.flatnonzero(comp.ravel() == q)
                if len(cells) < 8:
                    continue
                nodes = corners_of(g, cells)
                for t, c in enumerate(cells):
                    if int(c) in irr:
                        nodes[t] = d["irr_nodes"][irr[int(c)]]
                ref.append(np.unique(nodes.ravel()))
        assert len(got) == len(ref) == npieces, action
        for a, b in zip(got, ref):
            assert np.array_equal(a, b), action
