"""Drop-in for the engine preparation of ``rb/rbfe.py`` on the GPU: lattice material mixing
(``materials``, rb/rbfe.py:75-134), bone shares (``bone_share``, rb/replica3.py:17-29), the
bone part of every cell (rb/rbfe.py:140-141), and the preparation around the solve
(``prepare_solve``, rb/rbfe.py:146-178): ground contact, leg loads, the node sets of the bones,
the bone of every ligament end and the ligament patches; and ``solve`` (rb/rbfe.py:135-219), the
engine step around the split-node solve with its results: the region of every cell, the forces on
the talus and calcaneus from each region, the ground reactions by zone.

A lattice ``V`` (as built by the reference's voxeliser) holds ``L`` -- (ncell, ns^3) body labels
of the sub-samples of every cell in reference cell order (i ny + j) nz + k -- and ``cls`` -- the
tissue class of every body label.  ``materials`` returns the same (E, nu, active, F, ser) as the
reference; the per-cell work (class counts, the mixing rule, 3x3x3 joint detection) runs in
``cm_lattice_materials``.  There is no CPU path.

The class list, moduli, bone and cartilage body tables are the engine's definitions
(rb/rbgeom.py:35-49, rb/rbfe.py:29) and are kept identical.
"""
import ctypes
import json

import numpy as np
import torch

from ..system import _lib, require_gpu

# rb/rbgeom.py:35-49
COAL, FIB, INTERP, FUSE = 101, 102, 103, 104
GAP = 105
BONES = {"calcaneus": (1, 53), "talus": (2, 52), "fibula": (3, 54), "tibia": (4, 13), "navicular": (40, 45),
         "cuboid": (43, 49), "cun_med": (39, 51), "cun_int": (41, 46), "cun_lat": (42, 48)}
CART = {"ankle": 5, "syndesmosis": 6, "talofibular": 7, "mid": 8, "tn": 9, "ant": 10, "cc": 11, "post": 12}
CLASSES = ["void", "cortical", "cancellous", "cartilage", "fascia", "soft", "plate", "fibrous", "fat", "gap"]
MAT = {"cortical": (7300.0, 0.3), "cancellous": (100.0, 0.3), "cartilage": (10.0, 0.4), "fascia": (350.0, 0.4),
       "soft": (1.15, 0.49), "plate": (17000.0, 0.1), "fibrous": (60.0, 0.4), "fat": (1.15, 0.45),
       "gap": (0.01, 0.3)}
# rb/rbfe.py:29-30
ROLES = {"bones": dict(BONES), "cart": dict(CART), "bonded_extra": [COAL, FUSE]}
FORCES = {"tibia": 300.0, "fibula": 60.0}


def bone_groups(cls, roles):
    """rb/replica2.py:34-46: group id per bone body -- the named bones first (roles order), then
    every other bone body (cortical/cancellous, not bonded_extra) on its own in label order.
    Returns (group per body, name per group)."""
    grp, names = {}, {}
    nxt = 1
    for name, bodies in roles["bones"].items():
        for b in bodies:
            grp[b] = nxt
        names[nxt] = name
        nxt += 1
    bonded = set(roles.get("bonded_extra", ()))
    for b, c in sorted(cls.items()):
        if c in (1, 2) and b not in grp and b not in bonded:
            grp[b] = nxt
            names[nxt] = f"body{b}"
            nxt += 1
    return grp, names


def _class_lame(cart_E, nu_soft):
    mat = dict(MAT)
    mat["cartilage"] = (cart_E, 0.4)
    mat["soft"] = (1.15, nu_soft)
    E = np.array([1.0] + [mat[c][0] for c in CLASSES[1:]])
    nu = np.array([0.3] + [mat[c][1] for c in CLASSES[1:]])
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def lattice_prep(V, roles=ROLES, cart_E=10.0, nu_soft=0.49, device=None, host=True):
    """One GPU pass over the lattice: E, nu, active, F, ser (rb/rbfe.py:75-134), S and the bone
    names (rb/replica3.py:17-29), part (rb/rbfe.py:140-141).  Returns a dict of numpy arrays
    (``host``) or of device tensors."""
    dev = require_gpu(device)
    lib = _lib.load()
    cls = {int(k): int(v) for k, v in V["cls"].items()}
    L = np.asarray(V["L"])
    nx, ny, nz = len(V["xe"]) - 1, len(V["ye"]) - 1, len(V["ze"]) - 1
    ncell = nx * ny * nz
    if L.shape[0] != ncell:
        raise ValueError(f"L has {L.shape[0]} rows for {ncell} cells")
    ns3 = L.shape[1]
    nlab = max(max(cls), int(L.max())) + 1
    if nlab > 32767:
        raise ValueError("body labels must fit int16")
    lut_cls = np.zeros(nlab, np.int8)
    for b, c in cls.items():
        lut_cls[b] = c
    grp, names = bone_groups(cls, roles)
    ngrp = max(names) + 1 if names else 1
    lut_grp = np.zeros(nlab, np.int8)
    for b, g in grp.items():
        if b < nlab:
            lut_grp[b] = g
    lam_k, mu_k = _class_lame(cart_E, nu_soft)
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    Ld = t(L, torch.int16)
    lc, lg = t(lut_cls, torch.int8), t(lut_grp, torch.int8)
    lk, mk = t(lam_k, torch.float64), t(mu_k, torch.float64)
    E = torch.empty(ncell, dtype=torch.float64, device=dev)
    nu = torch.empty_like(E)
    act = torch.empty(ncell, dtype=torch.uint8, device=dev)
    ser = torch.empty_like(act)
    F = torch.empty((ncell, len(CLASSES)), dtype=torch.float64, device=dev)
    S = torch.empty((ncell, ngrp), dtype=torch.float32, device=dev)
    part = torch.empty(ncell, dtype=torch.int32, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.cm_lattice_materials(st, nx, ny, nz, ns3, vp(Ld), nlab, vp(lc), vp(lg), ngrp, vp(lk), vp(mk), vp(E),
                                        vp(nu), vp(act), vp(F), vp(ser), vp(S), vp(part)))
    out = dict(E=E, nu=nu, active=act.bool(), F=F, ser=ser.bool(), S=S, part=part, names=names)
    if host:
        out = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in out.items()}
    return out


def materials(V, roles=ROLES, cart_E=10.0, nu_soft=0.49, device=None):
    """rb/rbfe.py:75-134 on the GPU: (E, nu, active, F, ser)."""
    o = lattice_prep(V, roles, cart_E, nu_soft, device)
    return o["E"], o["nu"], o["active"], o["F"], o["ser"]


def bone_share(V, roles=ROLES, device=None):
    """rb/replica3.py:17-29 on the GPU: (S, names)."""
    o = lattice_prep(V, roles, device=device)
    return o["S"], o["names"]


# ---------------------------------------------------------------------------------------------
# Preparation around the solve (rb/rbfe.py:146-178)
def _dev(a, dt, dev):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)


def _vp(x):
    return ctypes.c_void_p(x.data_ptr())


def _geo(g, orig, dev):
    edges = [_dev(e, torch.float64, dev) for e in (g.xe, g.ye, g.ze)]
    od = _dev(orig, torch.int64, dev) if orig is not None else None
    return edges, od


def part_node_sets(cn, part, active, nnode, keys, device=None):
    """rb/rbfe.py:166-169,178 on the GPU: for every part id k in ``keys`` the sorted nodes of the
    active cells of part k, and the number of those cells.  Returns ({k: nodes}, {k: count})."""
    dev = require_gpu(device)
    lib = _lib.load()
    part = np.asarray(part)
    if len(keys) and (min(keys) < 1 or max(keys) > 63):
        raise ValueError("part ids must be 1..63")
    cnd = _dev(np.asarray(cn, np.int64), torch.int64, dev)
    pd = _dev(part, torch.int32, dev)
    ad = _dev(np.asarray(active, np.uint8), torch.uint8, dev)
    mask = torch.empty(int(nnode), dtype=torch.int64, device=dev)
    cnt = torch.empty(64, dtype=torch.int32, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.cm_part_nodes(st, len(part), _vp(cnd), _vp(pd), _vp(ad), int(nnode), _vp(mask), _vp(cnt)))
    m = mask.cpu().numpy().view(np.uint64)
    c = cnt.cpu().numpy()
    sets = {k: np.flatnonzero(m & np.uint64(1 << k)).astype(np.int64) for k in keys}
    return sets, {k: int(c[k]) for k in keys}


def _bone_lists(bone_nodes, names, dev):
    ptr = np.zeros(len(names) + 1, np.int64)
    for i, nm in enumerate(names):
        ptr[i + 1] = ptr[i] + len(bone_nodes[nm])
    flat = np.concatenate([np.asarray(bone_nodes[nm], np.int64) for nm in names]) if names else np.zeros(0, np.int64)
    return ptr, _dev(ptr, torch.int64, dev), _dev(flat if len(flat) else np.zeros(1, np.int64), torch.int64, dev)


def spring_end_bones(springs, bone_nodes, g, orig, device=None):
    """rb/rbfe.py:173-175 on the GPU: every spring end goes to the bone whose nearest node is
    nearest (first such bone in ``bone_nodes`` order; bones without nodes are skipped).  Node
    positions: their original's grid node (``orig``).  Returns [(bone of a, bone of b)]."""
    dev = require_gpu(device)
    lib = _lib.load()
    names = [nm for nm, nd in bone_nodes.items() if len(nd)]
    if not springs:
        return []
    if not names:
        raise ValueError("no bone has nodes")
    pts = np.array([p for s in springs for p in (s["a"], s["b"])], float).reshape(-1, 3)
    _, bptr, bnodes = _bone_lists(bone_nodes, names, dev)
    (xe, ye, ze), od = _geo(g, orig, dev)
    pd = _dev(pts, torch.float64, dev)
    d2 = torch.empty((len(pts), len(names)), dtype=torch.float64, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.cm_nearest_bone(st, len(pts), _vp(pd), len(names), _vp(bptr), _vp(bnodes), g.nx, g.ny, _vp(xe),
                                   _vp(ye), _vp(ze), _vp(od) if od is not None else None, _vp(d2)))
    d = np.sqrt(d2.cpu().numpy())
    best = d.argmin(1)          # first minimum: the first bone in order among equal distances
    return [(names[best[2 * i]], names[best[2 * i + 1]]) for i in range(len(springs))]


def spring_patches(springs, bone_nodes, g, orig, used=None, device=None):
    """The node patch of every ligament end (rb/replica3.py:74-80) on the GPU: the end's bone's
    used nodes within max(4 mm, distance of the 8th nearest).  ``springs`` carry their end bones in
    "A" and "B".  Returns [(patch of a, patch of b)] as sorted node arrays."""
    dev = require_gpu(device)
    lib = _lib.load()
    if not springs:
        return []
    names = list(bone_nodes)
    ptr, bptr, bnodes = _bone_lists(bone_nodes, names, dev)
    ends = [(p, names.index(s[key])) for s in springs for p, key in ((s["a"], "A"), (s["b"], "B"))]
    pts = np.array([p for p, _ in ends], float).reshape(-1, 3)
    bone = np.array([b for _, b in ends], np.int32)
    sizes = ptr[bone + 1] - ptr[bone]
    foff = np.r_[0, np.cumsum(sizes)].astype(np.int64)
    (xe, ye, ze), od = _geo(g, orig, dev)
    ud = _dev(np.asarray(used, np.uint8), torch.uint8, dev) if used is not None else None
    flags = torch.empty(max(int(foff[-1]), 1), dtype=torch.uint8, device=dev)
    rad = torch.empty(len(ends), dtype=torch.float64, device=dev)
    cnt = torch.empty(len(ends), dtype=torch.int32, device=dev)
    # device copies held in names until the kernel has run (a temporary's memory can be reused by
    # the next allocation before the launch)
    pd, bd, fd = _dev(pts, torch.float64, dev), _dev(bone, torch.int32, dev), _dev(foff, torch.int64, dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.cm_spring_patches(st, len(ends), _vp(pd), _vp(bd), _vp(bptr), _vp(bnodes),
                                     _vp(ud) if ud is not None else None, g.nx, g.ny, _vp(xe), _vp(ye), _vp(ze),
                                     _vp(od) if od is not None else None, _vp(fd), _vp(flags), _vp(rad), _vp(cnt)))
    fl = flags.cpu().numpy().astype(bool)
    if (cnt.cpu().numpy() == 0).any():
        raise ValueError("a spring end's bone has no used node")
    out = []
    for e, b in enumerate(bone):
        nodes = np.asarray(bone_nodes[names[b]], np.int64)
        out.append(nodes[fl[foff[e]:foff[e + 1]]])
    return [(out[2 * i], out[2 * i + 1]) for i in range(len(springs))]


def spring_matrix_cn(g, orig, springs, bone_nodes, device=None):
    """rb/replica3.py:68-86 with the patches found on the GPU: returns the reference's callback
    ``build(nmap, n)`` -> CSR k w w^T on compact DOFs, w = +d/n_A on the end-A patch and -d/n_B on
    the end-B patch (d the unit vector a -> b); candidates are the nodes with nmap >= 0.
    ``build.rank1(used)`` gives the same springs as rank-one terms (ptr, DOFs 3*node+c, w, k) over
    the full node layout, which the solver applies as k (w.u) w without forming w w^T (a 4 mm patch
    of a 0.5 mm grid holds ~2000 nodes: its dense block would have ~10^8 entries)."""
    import scipy.sparse as sp

    def weights(used):
        patches = spring_patches(springs, bone_nodes, g, orig, used=used, device=device)
        out = []
        for s, (pa, pb) in zip(springs, patches):
            a, b = np.asarray(s["a"], float), np.asarray(s["b"], float)
            d = (b - a) / np.linalg.norm(b - a)
            # per-DOF sums in the reference's order (end A's contribution first), nonzero entries
            # in DOF order
            dofs, ws = [], []
            for sel, sign in ((pa, 1.0), (pb, -1.0)):
                for c in range(3):
                    dofs.append(3 * np.asarray(sel, np.int64) + c)
                    ws.append(np.full(len(sel), sign * d[c] / len(sel)))
            dofs, ws = np.concatenate(dofs), np.concatenate(ws)
            ud, inv = np.unique(dofs, return_inverse=True)
            w = np.zeros(len(ud))
            np.add.at(w, inv, ws)
            keep = w != 0
            out.append((ud[keep], w[keep], float(s["k"])))
        return out

    def build(nmap, n):
        nmap = np.asarray(nmap, np.int64)
        rows, cols, vals = [], [], []
        for dofs, w, k in weights(nmap >= 0):
            nz = 3 * nmap[dofs // 3] + dofs % 3          # compact DOFs keep the full DOF order
            rows.append(np.repeat(nz, len(nz)))
            cols.append(np.tile(nz, len(nz)))
            vals.append(k * np.outer(w, w).ravel())
        if not rows:
            return sp.csr_matrix((n, n))
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))

    def rank1(used):
        terms = weights(np.asarray(used, bool))
        ptr = np.zeros(len(terms) + 1, np.int64)
        ptr[1:] = np.cumsum([len(t[0]) for t in terms])
        cat = lambda q, dt: np.concatenate([t[q] for t in terms]).astype(dt) if terms else np.zeros(0, dt)
        return ptr, cat(0, np.int64), cat(1, float), np.array([t[2] for t in terms], float)

    build.rank1 = rank1
    return build


def _face_nodes(g, cn, cells, up):
    """Four nodes of the z- (up False) or z+ face of every cell, tributary area = face / 4
    (rb/splitfe.py:94-102); corners l = a + 2b + 4c with c = up."""
    cells = np.asarray(cells, np.int64)
    cols = [4, 5, 6, 7] if up else [0, 1, 2, 3]
    i, j = cells // (g.ny * g.nz), (cells // g.nz) % g.ny
    area = g.dx[i] * g.dy[j] / 4.0
    return np.asarray(cn)[cells][:, cols].ravel(), np.repeat(area, 4)


def prepare_solve(V, g, cn, orig, active, part, names, roles=ROLES, forces=FORCES, load=1.0, rigid=True,
                  device=None):
    """rb/rbfe.py:146-178: everything ``solve_cn`` needs besides the materials.  ``names``: bone
    group id -> name (``lattice_prep``).  Returns dict(fixed=(nodes, area), loads=[(nodes, area*tz,
    dir)], bone_nodes={name: nodes}, springs=[...with "A"/"B"], extra=build callback or None,
    groups=[nodes per part with more than 3 active cells] or None).  The node sets, the spring ends
    and the ligament patches run on the GPU; the ground and top faces (one cell layer each) on the
    host."""
    active = np.asarray(active, bool)
    part = np.asarray(part)
    orig = np.asarray(orig, np.int64)
    cn = np.asarray(cn, np.int64)
    ncell = g.ncell
    kz = np.arange(ncell) % g.nz
    # ground contact: bottom-layer active cells under contact (fraction per (i, j) column)
    cf = np.asarray(V["contact"]).ravel()[np.arange(ncell) // g.nz]
    fc = np.flatnonzero((kz == 0) & active & (cf > 0))
    fn, fa = _face_nodes(g, cn, fc, up=False)
    fixed = (fn, fa * np.repeat(cf[fc], 4))
    # leg loads: top-layer active cells with more than half of the bone's sub-samples
    top = np.flatnonzero((kz == g.nz - 1) & active)
    Ltop = np.asarray(V["L"])[top]
    loads = []
    for name, Fz in forces.items():
        F = Fz * load
        frac = np.isin(Ltop, list(roles["bones"][name])).mean(1)
        n_, a_ = _face_nodes(g, cn, top[frac > 0.5], up=True)
        o = orig[n_]
        X = np.stack([g.xe[o % (g.nx + 1)], g.ye[(o // (g.nx + 1)) % (g.ny + 1)]], 1)
        W = a_.sum()
        c0 = (a_[:, None] * X).sum(0) / W
        dX = X - c0
        I = np.einsum("n,ni,nj->ij", a_, dX, dX)
        ab = np.linalg.solve(I, W * (np.asarray(V["load_line"][name], float) - c0))
        loads.append((n_, a_ * (-F / W * (1.0 + dX @ ab)), np.array([0.0, 0.0, 1.0])))
    # node sets of the bones and of every part (rigid groups)
    keys = sorted(names)
    sets, counts = part_node_sets(cn, part, active, len(orig), keys, device=device)
    bone_nodes = {}
    for bname in roles["bones"]:
        k = next(i for i, nm in names.items() if nm == bname)
        bone_nodes[bname] = sets[k]
    # ligament springs below the cut, each end on its nearest bone
    springs = [dict(s) for s in json.loads(V["springs_json"])]
    hc = float(np.diff(V["ze"]).max())
    springs = [s for s in springs if max(s["a"][2], s["b"][2]) < V["z_cut"] - 2 * hc]
    for s, (ea, eb) in zip(springs, spring_end_bones(springs, bone_nodes, g, orig, device=device)):
        s["A"], s["B"] = ea, eb
    extra = spring_matrix_cn(g, orig, springs, bone_nodes, device=device) if springs else None
    groups = [sets[k] for k in names if counts[k] > 3] if rigid else None
    return dict(fixed=fixed, loads=loads, bone_nodes=bone_nodes, springs=springs, extra=extra, groups=groups)


# ---------------------------------------------------------------------------------------------
# After the solve (rb/rbfe.py:184-219): stresses, region labels, forces on the talus and calcaneus,
# ground reactions by zone
REGIONS = list(CART) + ["coalition", "fibrous", "fat", "soft", "fascia", "bone", "other"]   # rb/rbfe.py:190


def _labels_dev(L, dev):
    if isinstance(L, torch.Tensor) and L.dtype == torch.int16 and L.device == dev:
        return L.contiguous()
    L = np.asarray(L)
    if L.size and (L.min() < 0 or L.max() > 32767):
        raise ValueError("body labels must fit int16")
    return torch.as_tensor(np.ascontiguousarray(L, np.int16), device=dev)


def body_fractions(L, sets, device=None):
    """bf(bodies) of rb/rbfe.py:138 (the fraction of a cell's sub-samples whose body is in the set)
    for up to 16 body sets in one GPU pass (cm_body_fractions), bit-exact to the numpy mean.
    L: (ncell, ns3) body labels, numpy or a device int16 tensor.  Returns a device FP64 tensor
    [len(sets)][ncell]."""
    dev = require_gpu(device)
    lib = _lib.load()
    if not 1 <= len(sets) <= 16:
        raise ValueError("1..16 body sets")
    Ld = _labels_dev(L, dev)
    ncell, ns3 = Ld.shape
    nlab = max([int(b) for bodies in sets for b in bodies] + [0]) + 1
    nlab = max(nlab, 32768 if ncell == 0 else int(Ld.max().item()) + 1)
    mask = np.zeros(nlab, np.uint32)
    for q, bodies in enumerate(sets):
        for b in bodies:
            if 0 <= int(b) < nlab:
                mask[int(b)] |= np.uint32(1 << q)
    md = torch.as_tensor(mask.view(np.int32), device=dev)
    out = torch.empty((len(sets), ncell), dtype=torch.float64, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.cm_body_fractions(st, ncell, ns3, _vp(Ld), nlab, _vp(md), len(sets), _vp(out)))
    return out


def region_labels(F, L, roles=ROLES, device=None):
    """rb/rbfe.py:190-201 on the GPU: region index (into ``REGIONS``-style names) of every cell --
    'other'; the dominant class of F (soft, fascia, or bone for cortical/cancellous); any
    coalition/fusion, fibrous, interposition body; any cartilage body, later rules winning.
    Returns (lab int16 numpy, names)."""
    dev = require_gpu(device)
    cart = roles["cart"]
    names = list(cart) + ["coalition", "fibrous", "fat", "soft", "fascia", "bone", "other"]
    sets = [(COAL, FUSE), (FIB,), (INTERP,)] + [(b,) for b in cart.values()]
    fr = body_fractions(L, sets, device=dev)
    Fd = torch.as_tensor(np.ascontiguousarray(F, np.float64), device=dev)
    dom = torch.argmax(Fd[:, 1:], 1) + 1          # first maximum, as numpy
    lab = torch.full((Fd.shape[0],), len(names) - 1, dtype=torch.int16, device=dev)

    def put(sel, name):
        lab.masked_fill_(sel, names.index(name))
    put(dom == CLASSES.index("soft"), "soft")
    put(dom == CLASSES.index("fascia"), "fascia")
    put((dom == 1) | (dom == 2), "bone")
    for q, name in enumerate(("coalition", "fibrous", "fat")):
        put(fr[q] > 0, name)
    for q, key in enumerate(cart):
        put(fr[3 + q] > 0, key)
    return lab.cpu().numpy(), names


def bone_forces(g, lam, mu, U, cn, part, active, k, lab, names, device=None):
    """rb/rbfe.py:202-210: the force on bone part k from each region -- minus the element forces
    K_e u_e (GPU, SRI) of the active cells outside the bone, on the bone's nodes, summed by the
    region of the cell.  Returns {name: [Fx, Fy, Fz]} unrounded, for the regions present."""
    from . import splitfe as SF
    cn = np.asarray(cn, np.int64)
    part, active, lab = np.asarray(part), np.asarray(active, bool), np.asarray(lab)
    nn = max(int(cn.max()) + 1 if cn.size else 0, len(U))
    tgt = (part == k) & active
    on_bone = np.zeros(nn, bool)
    on_bone[cn[tgt].ravel()] = True
    oc = np.flatnonzero(active & ~tgt)
    touch = on_bone[cn[oc]]
    oc = oc[touch.any(1)]
    fe = SF.element_forces_cn(g, lam, mu, U, oc, cn, device=device)
    fsum = -(fe * on_bone[cn[oc]][:, :, None]).sum(1)
    out = {}
    for i, nm in enumerate(names):
        sel = lab[oc] == i
        if sel.any():
            out[nm] = [float(v) for v in fsum[sel].sum(0)]
    return out


def ground_zones(R, x, scale, c0x):
    """rb/rbfe.py:211-215: vertical reactions summed over the heel (x < xh), mid, and fore (x >= xf)
    zones, xh = c0x + s (-40 - c0x), xf = c0x + s (20 - c0x) (unrounded)."""
    R, x = np.asarray(R, float), np.asarray(x, float)
    xh = c0x + scale * (-40.0 - c0x)
    xf = c0x + scale * (20.0 - c0x)
    heel, fore = x < xh, x >= xf
    return {"heel": float(R[heel, 2].sum()), "mid": float(R[~heel & ~fore, 2].sum()), "fore": float(R[fore, 2].sum())}


def solve(V, roles=ROLES, forces=FORCES, load=1.0, cart_E=10.0, tol=1e-6, rigid=True, nu_soft=0.49, keep=True,
          c0x=None, device=None, **opts):
    """rb/rbfe.py:135-219 on the GPU: lattice materials, bone parts, split connectivity, contact,
    loads and ligaments, the split-node solve, stresses, and the engine's results (res with the
    forces on the talus and calcaneus from every region and the ground reactions by zone).
    ``c0x``: the template's c0[0] (the zone edges of rb/rbfe.py:212-213 are in its frame; also read
    from V["c0x"]).  Returns the reference's dict (res, g, U, VM, SIG, part, bnames, cn, orig,
    active, F, lab, names, R, bone_nodes, E, nu)."""
    import time
    from ..grid import Grid, lame
    from . import splitfe as SF
    c0x = V.get("c0x", c0x) if hasattr(V, "get") else c0x
    if c0x is None:
        raise ValueError("solve needs the template's c0[0] (c0x) for the ground zones (rb/rbfe.py:211-215)")
    dev = require_gpu(device)
    t0 = time.time()
    g = Grid(V["xe"], V["ye"], V["ze"])
    o = lattice_prep(V, roles, cart_E, nu_soft, device=dev)
    E, nu, active, F, part, bnames = o["E"], o["nu"], o["active"], o["F"], o["part"], o["names"]
    lam, mu = lame(E, nu)
    cn, orig, nsplit = SF.split_connectivity_gpu(g, active, part, o["S"], device=dev)
    prep = prepare_solve(V, g, cn, orig, active, part, bnames, roles, forces, load, rigid, device=dev)
    t_prep = time.time() - t0
    U, info, R = SF.solve_cn(g, lam, mu, active, cn, orig, prep["fixed"], prep["loads"], prep["extra"],
                             prep["groups"], tol, device=dev, **opts)
    t1 = time.time()
    act = np.flatnonzero(active)
    sig, vm, _ = SF.cell_stress_cn(g, lam, mu, U, act, cn, device=dev)
    VM = np.zeros(g.ncell, np.float32)
    VM[act] = vm
    SIG = np.zeros((g.ncell, 6), np.float32)
    SIG[act] = sig
    res = {"ndof": int(info["ndof"]), "iters": info.get("iters"), "flag": info.get("flag"),
           "t_prep_s": round(t_prep, 1), "t_solve_s": round(t1 - t0 - t_prep, 1),
           "solver": {k: v for k, v in info.items() if k in ("t_hier", "t_gpu", "device", "relres", "coarse",
                                                           "gpu_attempt", "fallback", "gpu_error")},
           "n_split_nodes": int(nsplit), "contact_area_mm2": V.get("contact_area") if hasattr(V, "get") else None,
           "grid": list(g.shape)}
    lab, names = region_labels(F, V["L"], roles, device=dev)
    for bone in ("talus", "calcaneus"):
        k = next(i for i, nm in bnames.items() if nm == bone)
        f = bone_forces(g, lam, mu, U, cn, part, active, k, lab, names, device=dev)
        res[f"F_on_{bone}"] = {nm: [round(v, 2) for v in vec] for nm, vec in f.items()}
    xyz = g.node_xyz[orig]
    gz = ground_zones(R, xyz[:, 0], float(V["scale"]), float(c0x))
    res["ground_Fz"] = {k: round(v, 2) for k, v in gz.items()}
    res["ground_total_Fz"] = float(R[:, 2].sum())
    res["time_s"] = round(time.time() - t0, 1)
    return dict(res=res, g=g, U=U, VM=VM, SIG=SIG, part=part, bnames=bnames, cn=cn, orig=orig, active=active, F=F,
                lab=lab, names=names, R=R, bone_nodes=prep["bone_nodes"], E=E, nu=nu)


__all__ = ["COAL", "FIB", "INTERP", "FUSE", "GAP", "BONES", "CART", "CLASSES", "MAT", "ROLES", "FORCES", "REGIONS",
           "bone_groups", "lattice_prep", "materials", "bone_share", "part_node_sets", "spring_end_bones",
           "spring_patches", "spring_matrix_cn", "prepare_solve", "body_fractions", "region_labels", "bone_forces",
           "ground_zones", "solve"]
