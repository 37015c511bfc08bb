"""Seeded synthetic voxel rasters for the benchmark workloads (inputs in the reference's API).

``cube_inclusions``  32^3 cube at 1 mm, six random ellipsoidal inclusions (E 17000 in / 700 out,
    nu 0.3), bottom face held by stiff Winkler springs (k = 1e6 N/mm^3 x tributary area, the
    reference's "fixed" convention, rb/splitfe.py:143), 1 MPa compression on the top face.
    fe/solver.py ``solve`` form (2x2x2 integration).
``hindfoot_raster``  hindfoot-like raster (default 160x96x96 at 0.5 mm): soft-tissue envelope (E 1,
    nu 0.45) holding three ellipsoid bones (cortical shell 17000 / cancellous core 700, nu 0.3), a
    coalition cylinder (r = 4 voxels, E 700) bridging two of them, split nodes between bones
    (rb/splitfe.py:261-305) joined by isotropic springs k = 1e3 N/mm, ~2000 plantar ground springs
    k = 50 N/mm, 200 ligament springs k = 100 N/mm between bone-surface nodes, 700 N downward over
    the top 10 cell layers.  rb/splitfe.py ``solve_cn`` form (SRI).  The same generator at other
    sizes gives the uneven-grid batch and the large single solve (anatomy is relative to the box;
    shell thickness in voxels).
``batch_case``  64 feet x 8 surgical actions with grids between 96x64x64 and 192x112x112.

Seeds: ``numpy.random.default_rng(workload_id * 1000 + item)`` (cube 1, hindfoot 2, batch 3,
large 5).
"""
import numpy as np
from scipy import ndimage

from .grid import Grid, corners_of, face_node_areas, lame
from .rb.replica3 import cat_terms, spring_pairs, spring_terms
from .split import irregular_cells, split_connectivity_fast


def _rotation(rng, max_angle=None):
    if max_angle is None:
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(r))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        return q
    ang = rng.uniform(-max_angle, max_angle, 3)
    cx, cy, cz = np.cos(ang)
    sx, sy, sz = np.sin(ang)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _ellipsoid_radius(P, c, axes, R):
    """Normalised radius of points P (n,3) in an ellipsoid (centre c, semi-axes, rotation R)."""
    q = (P - c) @ R
    return np.sqrt(((q / axes) ** 2).sum(1))


def _cell_centres(g):
    return np.stack([g.xc[g.cI], g.yc[g.cJ], g.zc[g.cK]], 1)


# --------------------------------------------------------------------------------------------
def cube_inclusions(n=32, h=1.0, seed=1000, n_incl=6):
    rng = np.random.default_rng(seed)
    e = np.arange(n + 1) * h
    g = Grid(e, e, e)
    P = _cell_centres(g)
    inside = np.zeros(g.ncell, bool)
    for _ in range(n_incl):
        c = rng.uniform(0, n * h, 3)
        axes = rng.uniform(3, 8, 3) * h
        R = _rotation(rng)
        inside |= _ellipsoid_radius(P, c, axes, R) <= 1.0
    labels = inside.astype(np.uint8)                       # 0 matrix, 1 inclusion
    E_tab, nu_tab = np.array([700.0, 17000.0]), np.array([0.3, 0.3])
    lam_tab, mu_tab = lame(E_tab, nu_tab)
    lam, mu = lam_tab[labels], mu_tab[labels]
    active = np.ones(g.ncell, bool)
    bot = np.flatnonzero(g.cK == 0)
    top = np.flatnonzero(g.cK == g.nz - 1)
    nb, ab = face_node_areas(g, bot, "z-")
    nt, at = face_node_areas(g, top, "z+")
    springs = [(nb, ab, np.array([1e6, 1e6, 1e6]))]
    loads = [(nt, at, np.array([0.0, 0.0, -1.0]))]
    return dict(name="cube_inclusions", grid=g, lam=lam, mu=mu, active=active, labels=labels, lam_tab=lam_tab,
                mu_tab=mu_tab, springs=springs, loads=loads, sri=False)


def solve_inputs_cube(case):
    """cube_inclusions as level-0 GPU inputs: extra COO (diagonal springs) and nodal force vector."""
    g = case["grid"]
    nb, ab, kv = case["springs"][0]
    kd = np.zeros((g.nnode, 3))
    np.add.at(kd, nb, ab[:, None] * kv[None, :])
    nodes = np.flatnonzero(kd.any(1))
    rows = (3 * nodes[:, None] + np.arange(3)).ravel()
    extra = (rows, rows.copy(), kd[nodes].ravel())
    f = np.zeros((g.nnode, 3))
    nt, at, t = case["loads"][0]
    np.add.at(f, nt, at[:, None] * t[None, :])
    return extra, f.ravel()


# --------------------------------------------------------------------------------------------
FOOT_E = dict(soft=(1.0, 0.45), coalition=(700.0, 0.3), cortical=(17000.0, 0.3), cancellous=(700.0, 0.3))


def hindfoot_raster(dims=(160, 96, 96), h=0.5, seed=2000, shell=2, n_ground=2000, n_lig=200, k_iface=1e3, k_ground=50.0,
            k_lig=100.0, load=700.0, action="baseline", angle_seed=None):
    """Hindfoot-like raster.  action: baseline | resection | osteotomy | interposition | fusion |
    load0.5 | load1.5 | load2 (the surgical actions of the batched set)."""
    return hindfoot_family(dims, h, seed, shell, n_ground, n_lig, k_iface, k_ground, k_lig, load, [action],
                           angle_seed)[action]


# actions that keep the baseline's active cells and bone parts: they share its split connectivity,
# ground-spring nodes and ligament ends (the seeded draws depend on those only)
_SAME_TOPOLOGY = ("baseline", "resection", "fusion", "load0.5", "load1.5", "load2")


def hindfoot_family(dims=(160, 96, 96), h=0.5, seed=2000, shell=2, n_ground=2000, n_lig=200, k_iface=1e3,
                    k_ground=50.0, k_lig=100.0, load=700.0, actions=None, angle_seed=None):
    """``{action: hindfoot_raster(..., action=action)}`` for several actions of one foot, with the
    anatomy drawn once and the split connectivity, ground nodes and ligament ends computed once for
    all the actions that keep the baseline's cells (every action except osteotomy and
    interposition).  Same arrays as separate hindfoot_raster calls."""
    import copy
    actions = list(ACTIONS if actions is None else actions)
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    g = Grid(np.arange(nx + 1) * h, np.arange(ny + 1) * h, np.arange(nz + 1) * h)
    L = np.array([nx, ny, nz]) * h
    P = _cell_centres(g)
    env = _ellipsoid_radius(P, L * [0.5, 0.5, 0.45], L * [0.48, 0.45, 0.6], np.eye(3)) <= 1.0
    jit = lambda: rng.uniform(-0.03, 0.03, 3) * L
    bones = [  # calcaneus-, talus-, navicular-like
        (L * [0.30, 0.50, 0.33] + jit(), L * [0.20, 0.22, 0.20] * rng.uniform(0.95, 1.05, 3)),
        (L * [0.52, 0.50, 0.62] + jit(), L * [0.15, 0.20, 0.15] * rng.uniform(0.95, 1.05, 3)),
        (L * [0.70, 0.50, 0.47] + jit(), L * [0.12, 0.18, 0.15] * rng.uniform(0.95, 1.05, 3)),
    ]
    rots = [_rotation(rng, 0.25) for _ in bones]
    rad = np.stack([_ellipsoid_radius(P, c, a, R) for (c, a), R in zip(bones, rots)], 1)
    inb = (rad <= 1.0) & env[:, None]
    bone = np.where(inb.any(1), np.argmin(np.where(inb, rad, np.inf), 1) + 1, 0)
    del rad, inb
    # coalition: cylinder of radius 4 voxels bridging bones 1 and 2 (cells outside every bone)
    a0, a1 = bones[0][0], bones[1][0]
    d = a1 - a0
    t = np.clip(((P - a0) @ d) / (d @ d), 0, 1)
    dist = np.linalg.norm(P - (a0 + t[:, None] * d), axis=1)
    coal = (dist <= 4 * h) & env & (bone == 0)
    del t, dist
    # cortical shell: bone cells within `shell` voxels of the bone boundary
    shape = (nx, ny, nz)
    cort = np.zeros(g.ncell, bool)
    for b in (1, 2, 3):
        m = (bone == b).reshape(shape)
        core = ndimage.binary_erosion(m, iterations=shell)
        cort |= (m & ~core).ravel()
    labels0 = np.zeros(g.ncell, np.uint8)          # 0 soft, 1 coalition, 2+2(b-1) cortical, 3+2(b-1) cancellous
    labels0[coal] = 1
    isb = bone > 0
    labels0[isb] = (2 + 2 * (bone[isb] - 1) + (~cort[isb])).astype(np.uint8)
    E_lab = [FOOT_E["soft"], FOOT_E["coalition"]] + [FOOT_E["cortical"], FOOT_E["cancellous"]] * 3
    E_tab0 = np.array([e for e, _ in E_lab], float)
    nu_tab0 = np.array([v for _, v in E_lab], float)
    part0 = np.where(isb, bone, 0).astype(np.int64)
    rng_after_anatomy = copy.deepcopy(rng)
    slab = None
    if "osteotomy" in actions or "interposition" in actions:   # 3-voxel planar slab through bone 1
        arng = np.random.default_rng(seed if angle_seed is None else angle_seed)
        nrm = _rotation(arng, 0.5) @ np.array([1.0, 0.0, 0.0])
        slab = (np.abs((P - bones[0][0]) @ nrm) <= 1.5 * h) & (bone == 1)
    del P
    topo_cache = {}
    out = {}
    for action in actions:
        labels, active, part = labels0.copy(), env.copy(), part0.copy()
        E_tab, nu_tab = E_tab0, nu_tab0
        k_if, scale = k_iface, 1.0
        if action == "resection":                       # coalition relabelled to soft tissue
            labels[coal] = 0
        elif action in ("osteotomy", "interposition"):
            if action == "osteotomy":
                active &= ~slab
            else:                                       # gap filled with E = 5 MPa
                E_tab = np.append(E_tab, 5.0); nu_tab = np.append(nu_tab, 0.3)
                labels[slab] = len(E_tab) - 1
            part[slab] = 0
        elif action == "fusion":
            k_if = 1e6
        elif action.startswith("load"):
            scale = float(action[4:])
        elif action != "baseline":
            raise ValueError(f"unknown action {action!r}")
        labels[~active] = 255
        lam_tab, mu_tab = lame(E_tab, nu_tab)
        lab_safe = np.where(active, labels, 0)
        lam = np.where(active, lam_tab[lab_safe], 0.0)
        mu = np.where(active, mu_tab[lab_safe], 0.0)
        part[~active] = 0
        key = "base" if action in _SAME_TOPOLOGY else action
        if key not in topo_cache:
            topo_cache[key] = _topology(g, active, part, copy.deepcopy(rng_after_anatomy), n_ground, n_lig)
        tp = topo_cache[key]
        out[action] = _assemble_case(g, action, labels, active, part, lam, mu, lam_tab, mu_tab, tp, k_if, k_ground,
                                     k_lig, scale, load)
    return out


def _topology(g, active, part, rng, n_ground, n_lig):
    """Split connectivity, used nodes, ground-spring nodes and ligament ends of one foot's cells
    (the draws after the anatomy, in the generator's order)."""
    nx, ny, nz = g.nx, g.ny, g.nz
    share = np.zeros((g.ncell, 4), np.float32)
    share[np.arange(g.ncell), part] = 1.0
    share[:, 0] = 0.0
    cn, orig, nsplit = split_connectivity_fast(g, active, part, share)
    nn = len(orig)
    # nodes used by active cells
    used = np.zeros(nn, bool)
    used[cn[active].ravel()] = True
    xyz = g.node_xyz[orig]
    # plantar ground springs on used bottom grid nodes
    bottom = np.flatnonzero(used[: g.nnode] & (np.arange(g.nnode) // ((nx + 1) * (ny + 1)) == 0))
    gsel = np.sort(rng.choice(bottom, size=min(n_ground, len(bottom)), replace=False))
    # ligament springs between bone-surface nodes of different bones
    soft_touch = np.zeros(g.nnode, bool)
    soft_touch[corners_of(g, np.flatnonzero(active & (part == 0))).ravel()] = True
    surf = []
    for b in (1, 2, 3):
        nodes_b = np.unique(cn[np.flatnonzero(active & (part == b))].ravel())
        surf.append(nodes_b[soft_touch[orig[nodes_b]]])
    pairs = []
    for _ in range(n_lig):
        b1, b2 = rng.choice(3, 2, replace=False)
        if len(surf[b1]) == 0 or len(surf[b2]) == 0:
            continue
        a = rng.choice(surf[b1])
        dd = np.linalg.norm(xyz[surf[b2]] - xyz[a], axis=1)
        near = surf[b2][np.argsort(dd)[:50]]
        b_ = rng.choice(near)
        if np.linalg.norm(xyz[b_] - xyz[a]) > 0:
            pairs.append((a, b_))
    return dict(share=share, cn=cn, orig=orig, nsplit=nsplit, used=used, xyz=xyz, gsel=gsel,
                pairs=np.asarray(pairs, np.int64).reshape(-1, 2))


def _assemble_case(g, action, labels, active, part, lam, mu, lam_tab, mu_tab, tp, k_if, k_ground, k_lig, scale,
                   load):
    """One action's case dict from its materials, the foot's topology and the spring/load values."""
    cn, orig, xyz, pairs = tp["cn"], tp["orig"], tp["xyz"], tp["pairs"]
    nn = len(orig)
    # interface springs: copy <-> its original grid node
    copies = np.arange(g.nnode, nn)
    cpairs = np.stack([copies, orig[copies]], 1)
    iface = spring_pairs(xyz, cpairs, k_if, directional=False)
    lig = spring_pairs(xyz, pairs, k_lig, directional=True)
    terms = cat_terms(spring_terms(xyz, cpairs, k_if, directional=False),
                      spring_terms(xyz, pairs, k_lig, directional=True))
    # load: downward total over active cells of the top 10 layers, 1/8 per corner
    topc = np.flatnonzero(active & (g.cK >= g.nz - 10))
    ln = cn[topc].ravel()
    la = np.full(len(ln), scale * load / len(ln))
    loads = [(ln, la, np.array([0.0, 0.0, -1.0]))]
    gsel = tp["gsel"]
    fixed = (gsel, np.full(len(gsel), k_ground / 1e6))
    extra = (np.concatenate([iface[0], lig[0]]), np.concatenate([iface[1], lig[1]]),
             np.concatenate([iface[2], lig[2]]))

    def extra_cb(nmap, n):
        import scipy.sparse as sp
        r, c, v = extra
        ok = (nmap[r // 3] >= 0) & (nmap[c // 3] >= 0)
        rr = 3 * nmap[r[ok] // 3] + r[ok] % 3
        cc = 3 * nmap[c[ok] // 3] + c[ok] % 3
        return sp.csr_matrix((v[ok], (rr, cc)), shape=(n, n))

    return dict(name="hindfoot_raster", grid=g, lam=lam, mu=mu, active=active, labels=labels, lam_tab=lam_tab,
                mu_tab=mu_tab, part=part, bone_share=tp["share"], cn=cn, orig=orig, nsplit=tp["nsplit"], fixed=fixed,
                loads=loads, extra=extra, extra_cb=extra_cb, terms=terms, sri=True, action=action, used=tp["used"],
                n_ground=len(gsel), n_lig=len(pairs), n_copies=len(copies))


def solve_inputs_hindfoot(case, as_terms=True):
    """hindfoot_raster as level-0 GPU inputs: extra COO over node DOFs (ground springs; with
    ``as_terms=False`` also the interface and ligament springs as COO), rank-one terms (interface
    and ligament springs) and the nodal force vector (grid nodes then copies)."""
    nn = len(case["orig"])
    gs, ga = case["fixed"]
    rows = (3 * np.asarray(gs)[:, None] + np.arange(3)).ravel()
    ground = (rows, rows.copy(), np.repeat(ga * 1e6, 3))
    f = np.zeros((nn, 3))
    for nodes, area, t in case["loads"]:
        np.add.at(f, nodes, area[:, None] * t[None, :])
    if as_terms:
        return ground, case["terms"], f.ravel()
    r, c, v = case["extra"]
    extra = (np.concatenate([ground[0], r]), np.concatenate([ground[1], c]), np.concatenate([ground[2], v]))
    return extra, None, f.ravel()


ACTIONS = ["baseline", "resection", "osteotomy", "interposition", "fusion", "load0.5", "load1.5", "load2"]


def batch_sizes(n_feet=64, seed=3000):
    """Grid sizes of the batched configuration: uniform between 96x64x64 and 192x112x112 (even)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.array([96, 64, 64]), np.array([192, 112, 112])
    return [tuple(int(2 * round(v / 2)) for v in rng.uniform(lo, hi)) for _ in range(n_feet)]


def batch_case(foot, action, n_feet=64, h=0.5, scale=1):
    """Foot ``foot`` of the batched set under ``action``.  ``scale`` > 1 draws the same foot on
    coarser voxels (each grid dimension / scale, rounded to even, voxel size h * scale): the same
    anatomy at a size the FP64 oracle solves in seconds."""
    dims = batch_sizes(n_feet)[foot]
    if scale != 1:
        dims = tuple(max(4, int(2 * round(d / scale / 2))) for d in dims)
    return hindfoot_raster(dims=dims, h=h * scale, seed=3000 + foot, action=action, angle_seed=3500 + foot)


def hindfoot_large(dims=(384, 256, 256), h=0.25, seed=5000):
    return hindfoot_raster(dims=dims, h=h, seed=seed, shell=4)


def level0_description(case):
    """The GPU solver's inputs of a hindfoot case in compact form (numpy only): grid edges, labels
    and their (lambda, mu) tables (255 = inactive), the bone part of every cell (0 = none), split
    copies and the cells touching them, ground springs as COO, interface and ligament springs as
    rank-one terms, and the non-zero nodal loads (the load acts on a few top layers only)."""
    g = case["grid"]
    extra, terms, f = solve_inputs_hindfoot(case)
    irr_c, irr_n = irregular_cells(g, case["active"], case["cn"])
    nz = np.flatnonzero(f)
    return dict(edges=(g.xe, g.ye, g.ze), labels=case["labels"], part=np.asarray(case["part"]).astype(np.uint8),
                lam_tab=case["lam_tab"], mu_tab=case["mu_tab"],
                copy_orig=case["orig"][g.nnode:], irr_cells=irr_c, irr_nodes=irr_n, extra=extra, terms=terms,
                f_idx=nz, f_val=f[nz], nn=len(case["orig"]), action=case["action"], ncell=g.ncell,
                nactive=int(case["active"].sum()))


def batch_foot(foot, n_feet=64, h=0.5, scale=1, actions=None):
    """level0_description of one foot of the batched set under its surgical actions (all 8 by
    default, in ACTIONS order); the anatomy and the shared topology are computed once."""
    dims = batch_sizes(n_feet)[foot]
    if scale != 1:
        dims = tuple(max(4, int(2 * round(d / scale / 2))) for d in dims)
    fam = hindfoot_family(dims, h * scale, 3000 + foot, actions=actions, angle_seed=3500 + foot)
    return [level0_description(fam[a]) for a in (ACTIONS if actions is None else actions)]
