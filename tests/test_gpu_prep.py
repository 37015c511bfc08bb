"""GPU engine preparation against the oracle: lattice material mixing, bone shares and part
(rb/rbfe.py:75-141, rb/replica3.py:17-29) and split connectivity (rb/splitfe.py:261-305).
Booleans, fractions, shares, parts and node numbering bit-exact; (E, nu) to 1e-12 relative
(FP64, the same formulas evaluated in the same order)."""
import numpy as np
import pytest

from comech_amd.grid import Grid
from comech_amd.rb import rbfe as RF
from comech_amd.rb import splitfe as SF
from oracle import fe_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

# body labels -> classes: bones (cortical 1 / cancellous 2), cartilage bodies (3), fascia (4), soft
# (5), plate (6), coalition bars (cortical, bonded to the bones they touch), fibrous (7), fat (8),
# gap (9), and one unnamed bone body (its own group)
CLS = {1: 1, 53: 2, 2: 1, 52: 2, 3: 1, 54: 2, 40: 1, 45: 2, 5: 3, 8: 3, 12: 3, 60: 4, 61: 5, 62: 6,
       RF.COAL: 1, RF.FIB: 7, RF.INTERP: 8, RF.FUSE: 1, RF.GAP: 9, 77: 2}
LABELS = [0, 0, 1, 53, 2, 52, 3, 54, 40, 45, 5, 8, 12, 60, 61, 61, 62, RF.COAL, RF.FIB, RF.INTERP, RF.FUSE, RF.GAP, 77]


def _lattice(dims, ns3, seed, blobs=True):
    rng = np.random.default_rng(seed)
    nc = int(np.prod(dims))
    L = rng.choice(LABELS, size=(nc, ns3))
    if blobs:   # spatially coherent bone regions so that joints and parts occur
        nx, ny, nz = dims
        i, j, k = np.unravel_index(np.arange(nc), dims)
        L[(i < nx // 2) & (rng.uniform(size=nc) < 0.8)] = 1
        L[(i >= nx // 2) & (j < ny // 2) & (rng.uniform(size=nc) < 0.8)] = 52
    L[rng.uniform(size=nc) < 0.1] = 0            # empty cells
    return dict(L=L, cls=CLS, xe=np.arange(dims[0] + 1.0), ye=np.arange(dims[1] + 1.0), ze=np.arange(dims[2] + 1.0))


@pytest.mark.parametrize("seed,ns3,blobs", [(0, 27, True), (1, 27, False), (2, 8, True), (3, 8, False)])
def test_lattice_materials_match_oracle(seed, ns3, blobs):
    dims = (9, 7, 6)
    V = _lattice(dims, ns3, seed, blobs)
    o = RF.lattice_prep(V, cart_E=12.0, nu_soft=0.48, device=DEV)
    E, nu, act, F, ser = O.lattice_materials(V["L"], CLS, RF.ROLES, dims, cart_E=12.0, nu_soft=0.48)
    S, names = O.bone_share(V["L"], CLS, RF.ROLES)
    part = O.lattice_part(S, F)
    assert np.array_equal(o["active"], act) and np.array_equal(o["ser"], ser)
    assert np.array_equal(o["F"], F) and np.array_equal(o["S"], S) and o["names"] == names
    assert np.array_equal(o["part"], part)
    assert np.allclose(o["E"], E, rtol=1e-12, atol=0) and np.allclose(o["nu"], nu, rtol=1e-12, atol=0)
    assert act.any() and ser.any() and (~act).any()
    assert (part > 0).any() or not blobs          # random labels alone rarely fill half a cell with one bone


@pytest.mark.parametrize("seed", range(4))
def test_split_connectivity_gpu_bit_exact(seed):
    """Random parts, activity and shares (ties and zero-share cells included)."""
    rng = np.random.default_rng(200 + seed)
    shape = (9, 7, 8)
    g, og = Grid(*(np.arange(n + 1.0) for n in shape)), O.Grid(*(np.arange(n + 1.0) for n in shape))
    nb = 5
    part = rng.integers(0, nb, g.ncell)
    active = rng.uniform(size=g.ncell) > 0.1
    share = rng.uniform(size=(g.ncell, nb)).astype(np.float32)
    share[rng.uniform(size=g.ncell) < 0.3] = 0
    share[:, 0] = 0
    share[rng.uniform(size=g.ncell) < 0.2, 2] = share[rng.uniform(size=g.ncell) < 1.0, 1][0]   # ties
    cn, orig, ns = SF.split_connectivity_gpu(g, active, part, share, device=DEV)
    cno, origo, nso = O.split_connectivity(og, active, part, share)
    assert ns == nso and np.array_equal(orig, origo) and np.array_equal(cn, cno)


def test_split_connectivity_gpu_hindfoot_raster():
    """The synthetic hindfoot raster's bones: GPU == the product's host form == oracle."""
    from comech_amd import synth
    c = synth.hindfoot_raster(dims=(40, 24, 24), h=2.0)
    g = c["grid"]
    cn, orig, ns = SF.split_connectivity_gpu(g, c["active"], c["part"], c["bone_share"], device=DEV)
    cnh, origh, nsh = SF.split_connectivity_fast(g, c["active"], c["part"], c["bone_share"])
    assert ns == nsh > 0 and np.array_equal(orig, origh) and np.array_equal(cn, cnh)


def test_lattice_then_split_chain():
    """Lattice prep (shares, part, activity) feeding the GPU split connectivity, as rb/rbfe.py:139-142
    chains them, equals the oracle chain."""
    dims = (10, 8, 7)
    V = _lattice(dims, 27, 9)
    o = RF.lattice_prep(V, device=DEV)
    g, og = Grid(*(np.arange(n + 1.0) for n in dims)), O.Grid(*(np.arange(n + 1.0) for n in dims))
    cn, orig, ns = SF.split_connectivity_gpu(g, o["active"], o["part"], o["S"], device=DEV)
    E, nu, act, F, ser = O.lattice_materials(V["L"], CLS, RF.ROLES, dims)
    S, _ = O.bone_share(V["L"], CLS, RF.ROLES)
    cno, origo, nso = O.split_connectivity(og, act, O.lattice_part(S, F), S)
    assert ns == nso and np.array_equal(orig, origo) and np.array_equal(cn, cno)


# ---------------------------------------------------------------------------------------------
# Preparation around the solve (rb/rbfe.py:146-178)
CLS_E = dict(CLS)
CLS_E.update({4: 1, 13: 2})


def _engine_V(dims, seed):
    """A lattice with calcaneus/talus below, tibia and fibula in the top layers, voids, a ground
    contact map, load lines off the loaded areas' centroids and ligament springs."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    nc = nx * ny * nz
    i, j, k = np.unravel_index(np.arange(nc), dims)
    L = rng.choice([0, 61, 60, 5], size=(nc, 27))
    for lab, sel in ((1, k < nz // 3), (52, (k >= nz // 3) & (k < 2 * nz // 3)), (4, (k >= 2 * nz // 3) & (i < nx // 2)),
                     (3, (k >= 2 * nz // 3) & (i >= nx // 2))):
        L[sel & (rng.uniform(size=nc) < 0.9)] = lab
    L[rng.uniform(size=nc) < 0.05] = 0
    xe, ye, ze = np.arange(nx + 1.0), np.arange(ny + 1.0), np.arange(nz + 1.0)
    contact = (rng.uniform(size=(nx, ny)) * (rng.uniform(size=(nx, ny)) > 0.3)).astype(np.float32)
    springs = [{"name": "", "k": float(rng.uniform(50, 150)), "a": [float(rng.uniform(0, nx)), float(rng.uniform(0, ny)),
                                                                   float(rng.uniform(0, nz - 3))],
                "b": [float(rng.uniform(0, nx)), float(rng.uniform(0, ny)), float(rng.uniform(0, nz - 3))]}
               for _ in range(12)]
    springs.append({"name": "", "k": 1.0, "a": [1.0, 1.0, nz - 0.5], "b": [2.0, 2.0, nz - 0.2]})   # above the cut
    import json
    return dict(L=L, cls=CLS_E, xe=xe, ye=ye, ze=ze, contact=contact,
                load_line={"tibia": [nx / 4 + 0.3, ny / 2 - 0.2], "fibula": [3 * nx / 4 - 0.4, ny / 2 + 0.1]},
                springs_json=json.dumps(springs), z_cut=float(nz))


@pytest.mark.parametrize("seed", [0, 1])
def test_prepare_solve_matches_oracle(seed):
    dims = (14, 9, 12)
    V = _engine_V(dims, seed)
    o = RF.lattice_prep(V, device=DEV)
    g, og = Grid(V["xe"], V["ye"], V["ze"]), O.Grid(V["xe"], V["ye"], V["ze"])
    cn, orig, ns = SF.split_connectivity_gpu(g, o["active"], o["part"], o["S"], device=DEV)
    P = RF.prepare_solve(V, g, cn, orig, o["active"], o["part"], o["names"], device=DEV)
    act, part, names = o["active"], o["part"], o["names"]
    # ground contact and loads
    fn, fa = O.ground_contact(og, cn, act, V["contact"])
    assert np.array_equal(P["fixed"][0], fn) and np.allclose(P["fixed"][1], fa, rtol=1e-15, atol=0)
    frac = {nm: np.isin(V["L"], list(RF.ROLES["bones"][nm])).mean(1) for nm in RF.FORCES}
    loads = O.leg_loads(og, cn, orig, act, frac, RF.FORCES, 1.0, V["load_line"])
    assert len(loads) == len(P["loads"]) == 2
    for (n1, f1, d1), (n2, f2, d2) in zip(P["loads"], loads):
        assert np.array_equal(n1, n2) and np.array_equal(d1, d2) and len(n1) > 0
        assert np.allclose(f1, f2, rtol=1e-12, atol=1e-12 * np.abs(f2).max())
    # node sets and rigid groups
    keys = sorted(names)
    sets = O.bone_node_sets(cn, part, act, keys)
    counts = O.part_cell_counts(part, act, keys)
    for bname, nodes in P["bone_nodes"].items():
        k = next(i for i, nm in names.items() if nm == bname)
        assert np.array_equal(nodes, sets[k]), bname
    ref_groups = [sets[k] for k in names if counts[k] > 3]
    assert len(P["groups"]) == len(ref_groups) >= 4
    assert all(np.array_equal(a, b) for a, b in zip(P["groups"], ref_groups))
    # spring ends (one spring is above the cut) and the ligament stiffness
    xyz = og.node_xyz[orig]
    assert len(P["springs"]) == 12
    ends = O.spring_end_bones(P["springs"], P["bone_nodes"], xyz)
    assert [(s["A"], s["B"]) for s in P["springs"]] == ends
    rng = np.random.default_rng(seed)
    used = rng.uniform(size=len(orig)) > 0.15                 # a used map that drops some bone nodes
    nmap = -np.ones(len(orig), np.int64)
    nmap[used] = np.arange(used.sum())
    n = 3 * int(used.sum())
    Kg = P["extra"](nmap, n).tocsr()
    Ko = O.spring_matrix_cn(xyz, P["springs"], P["bone_nodes"])(nmap, n).tocsr()
    assert Kg.nnz == Ko.nnz and np.array_equal(Kg.indices, Ko.indices) and np.array_equal(Kg.indptr, Ko.indptr)
    assert np.abs(Kg.data - Ko.data).max() <= 1e-13 * np.abs(Ko.data).max()


def test_spring_patches_radius_rule():
    """The patch is the 8 nearest used nodes when they lie beyond 4 mm, every used node within 4 mm
    otherwise, all candidates when fewer than 8; checked against a sort of the distances."""
    rng = np.random.default_rng(11)
    g = Grid(*(np.arange(n + 1.0) * 0.9 for n in (30, 20, 18)))
    og = O.Grid(g.xe, g.ye, g.ze)
    nodes = np.sort(rng.choice(g.nnode, 5000, replace=False))
    bone_nodes = {"x": nodes, "few": nodes[:5]}
    used = rng.uniform(size=g.nnode) > 0.1
    springs = []
    for p in ([1.0, 1.0, 1.0], [13.0, 9.0, 8.0], [40.0, 40.0, 40.0], [5.5, 3.3, 2.2]):
        springs.append({"a": p, "b": [0.0, 0.0, 0.0], "A": "x", "B": "few", "k": 1.0})
    patches = RF.spring_patches(springs, bone_nodes, g, None, used=used, device=DEV)
    for s, (pa, pb) in zip(springs, patches):
        for p, key, pp in ((s["a"], "A", pa), (s["b"], "B", pb)):
            cand = bone_nodes[s[key]][used[bone_nodes[s[key]]]]
            d = np.linalg.norm(og.node_xyz[cand] - np.asarray(p), axis=1)
            r = max(4.0, np.sort(d)[min(7, len(d) - 1)])
            assert np.array_equal(pp, cand[d <= r])
            assert len(pp) >= min(8, len(cand))


def test_ligament_rank1_terms_equal_matrix():
    """build.rank1(used) -- the springs as rank-one terms over node DOFs -- gives the same k w w^T as
    the CSR callback on compact DOFs."""
    V = _engine_V((12, 9, 10), 5)
    o = RF.lattice_prep(V, device=DEV)
    g = Grid(V["xe"], V["ye"], V["ze"])
    cn, orig, ns = SF.split_connectivity_gpu(g, o["active"], o["part"], o["S"], device=DEV)
    P = RF.prepare_solve(V, g, cn, orig, o["active"], o["part"], o["names"], device=DEV)
    used = np.random.default_rng(2).uniform(size=len(orig)) > 0.1
    nmap = -np.ones(len(orig), np.int64)
    nmap[used] = np.arange(used.sum())
    n = 3 * int(used.sum())
    K = P["extra"](nmap, n).toarray()
    ptr, dofs, w, k = P["extra"].rank1(used)
    assert len(k) == len(P["springs"]) and ptr[-1] == len(dofs) == len(w)
    R = np.zeros((n, n))
    for t in range(len(k)):
        d = dofs[ptr[t]:ptr[t + 1]]
        cdof = 3 * nmap[d // 3] + d % 3
        assert (nmap[d // 3] >= 0).all()
        R[np.ix_(cdof, cdof)] += k[t] * np.outer(w[ptr[t]:ptr[t + 1]], w[ptr[t]:ptr[t + 1]])
    assert np.abs(R - K).max() <= 1e-13 * np.abs(K).max()
