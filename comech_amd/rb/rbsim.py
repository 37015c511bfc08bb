"""Drop-ins for ``sample_fe`` and ``sample_nodes`` of the reference's ``rb/rbsim.py:110-161`` on the GPU,
and for its effect descriptors (``rigid_fit`` and ``effects``, rb/rbsim.py:99-107,164-203).

sample_fe: every analysis cell averages von Mises stress, stress and displacement over sub^3 sample
points in the FE cells; samples outside the FE box or in inactive FE cells are excluded.  The host
builds the per-axis sample tables (cell index, local coordinate, inside flag); the averaging runs in
the ``k_sample_fe`` kernel.  Returns (vm, u, cover) and sets ``sample_fe.sig`` like the reference.

sample_nodes: the relative displacement at every analysis-grid node, trilinear in the FE cell that
contains the node, or in the nearest active FE cell when that one is empty (the feature transform
of scipy.ndimage.distance_transform_edt in cell-index units, rb/rbsim.py:152, computed by the
``k_column_nearest`` / ``k_sample_nodes`` kernels with scipy's tie rule: smallest (k, j, i)).
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from ..system import require_gpu


def _dev(a, dtype, dev):
    """numpy array or torch tensor -> contiguous flat device tensor of ``dtype`` (no copy when it
    already is one: the engine's displacement and connectivity can stay on the GPU)."""
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype).contiguous().reshape(-1)
    np_dt = torch.empty((), dtype=dtype).numpy().dtype
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, np_dt).reshape(-1)), device=dev)


def _axis(edges, d, ml_edges, ml_d, sub):
    oo = (np.arange(sub) + 0.5) / sub
    X = (ml_edges[:-1, None] + ml_d[:, None] * oo).ravel()
    n = len(edges) - 1
    idx = np.clip(np.searchsorted(edges, X) - 1, 0, n - 1)
    t = np.clip((X - edges[idx]) / d[idx], 0, 1)
    inside = (X >= edges[0]) & (X <= edges[-1])
    return idx.astype(np.int32), t, inside.astype(np.uint8)


def sample_fe(o, grid_ml, Urel, sub=2, device=None):
    dev = require_gpu(device)
    lib = _lib.load()
    g, cn, act = o["g"], o["cn"], o["active"]
    ax = [_axis(e, d, me, md, sub) for e, d, me, md in
          ((g.xe, g.dx, grid_ml.xe, grid_ml.dx), (g.ye, g.dy, grid_ml.ye, grid_ml.dy),
           (g.ze, g.dz, grid_ml.ze, grid_ml.dz))]
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    tabs = []
    for idx, tt, ins in ax:
        tabs += [t(idx, torch.int32), t(tt, torch.float64), t(ins, torch.uint8)]
    mx, my, mz = grid_ml.shape
    ncm = mx * my * mz
    a = _dev(act, torch.uint8, dev)
    c = _dev(cn, torch.int64, dev)
    vm_in = _dev(o["VM"], torch.float32, dev)
    sg_in = _dev(o["SIG"], torch.float32, dev)
    U = _dev(Urel, torch.float64, dev)
    vm = torch.empty(ncm, dtype=torch.float32, device=dev)
    u = torch.empty(ncm * 3, dtype=torch.float32, device=dev)
    cov = torch.empty(ncm, dtype=torch.float64, device=dev)
    sg = torch.empty(ncm * 6, dtype=torch.float32, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    _lib.check(lib.cm_sample_fe(ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream), mx, my, mz, sub, g.nx, g.ny,
                                g.nz, *[vp(x) for x in tabs], vp(a), vp(c), vp(vm_in), vp(sg_in), vp(U), vp(vm), vp(u),
                                vp(cov), vp(sg)))
    sample_fe.sig = sg.view(ncm, 6).cpu().numpy()
    return vm.cpu().numpy(), u.view(ncm, 3).cpu().numpy(), cov.cpu().numpy()


def sample_nodes(o, grid_ml, Urel, device=None, noalloc=False):
    """rb/rbsim.py:143-161.  o: dict with the FE grid ``g``, connectivity ``cn`` and ``active``;
    grid_ml: the analysis grid; Urel: (nodes, 3) relative displacement (NaN for unused nodes).
    cn, active and Urel may be numpy arrays or device tensors.  Returns float32 (nnode_ml, 3) in
    the analysis grid's node order.  ``noalloc``: run the library's allocation-free entry point
    (its work tables in the library's per-device scratch) instead of passing work arrays."""
    dev = require_gpu(device)
    lib = _lib.load()
    g, cn = o["g"], o["cn"]
    a = _dev(o["active"], torch.uint8, dev)
    if a.numel() != g.nx * g.ny * g.nz:
        raise ValueError("active must hold one flag per FE cell")
    if not bool(a.any()):
        raise ValueError("sample_nodes needs at least one active FE cell")
    cells = []
    for edges, n, me in ((g.xe, g.nx, grid_ml.xe), (g.ye, g.ny, grid_ml.ye), (g.ze, g.nz, grid_ml.ze)):
        cells.append(np.clip(np.searchsorted(edges, me) - 1, 0, n - 1).astype(np.int32))
    kq, qz = np.unique(cells[2], return_inverse=True)
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    me = [t(np.asarray(e, np.float64), torch.float64) for e in (grid_ml.xe, grid_ml.ye, grid_ml.ze)]
    ce = [t(c, torch.int32) for c in cells]
    kqd, qzd = t(kq.astype(np.int32), torch.int32), t(qz.astype(np.int32), torch.int32)
    fe = [t(np.asarray(e, np.float64), torch.float64) for e in (g.xe, g.ye, g.ze)]
    c = _dev(cn, torch.int64, dev)
    U = _dev(Urel, torch.float64, dev)
    mx, my, mz = len(grid_ml.xe) - 1, len(grid_ml.ye) - 1, len(grid_ml.ze) - 1
    out = torch.empty(((mx + 1) * (my + 1) * (mz + 1), 3), dtype=torch.float32, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    args = (mx, my, mz, *[vp(x) for x in me], *[vp(x) for x in ce], len(kq), vp(kqd), vp(qzd), g.nx, g.ny, g.nz,
            *[vp(x) for x in fe], vp(a), vp(c), vp(U))
    with torch.cuda.device(dev):
        if noalloc:
            _lib.check(lib.cm_sample_nodes_noalloc(st, *args, vp(out)))
        else:
            work = torch.empty((2, len(kq) * g.nx * g.ny), dtype=torch.int32, device=dev)
            _lib.check(lib.cm_sample_nodes(st, *args, vp(work[0]), vp(work[1]), vp(out)))
    return out.cpu().numpy()


def rigid_fit(X, U):
    """rb/rbsim.py:99-107: the least-squares rigid motion t + w x (X - c) of the points X (c their
    centroid) closest to the displacements U.  About the centroid the translation and the rotation
    decouple: t = mean(U), and w solves J w = sum_i r_i x u_i with J = sum_i (|r_i|^2 I - r_i r_i^T)
    (the normal equations of the reference's 6-column system).  Returns (t, w, c)."""
    X = np.asarray(X, float)
    U = np.asarray(U, float)
    c = X.mean(0)
    r = X - c
    t = U.mean(0)
    J = np.eye(3) * (r * r).sum() - r.T @ r
    w = np.linalg.solve(J, np.cross(r, U - t).sum(0))
    return t, w, c


def effects(o, foot=None, act=None, T=None, dt=None, dc=None, device=None):
    """rb/rbsim.py:164-203: the engine's effect descriptors from a solve (the dict of
    comech_amd.rb.rbfe.solve with o["L"] = the lattice labels).  Cartilage p99 and volume-weighted
    mean von Mises (ant, mid, post, tn, cc), the talus load shares, the talus translation and eversion
    relative to the calcaneus, the navicular drop, the heel share, the coalition p99 and max; tc_p99
    (the talocalcaneal band, rb/rbsim.py:193-199) needs the template's distances of the cell centres
    to the talus and calcaneus (``dt``, ``dc``): computed when they are given, left out otherwise
    (foot, act, T are the reference's template arguments and are not used).  The body fractions
    run on the GPU (cm_body_fractions).  Returns (d, Urel)."""
    from . import rbfe as RF
    g = o["g"]
    VM = np.asarray(o["VM"])
    active = np.asarray(o["active"], bool)
    part = np.asarray(o["part"])
    U = np.asarray(o["U"], float)
    vol = g.cell_vol
    cart = RF.CART
    keys = ("ant", "mid", "post", "tn", "cc")
    sets = [(cart[k],) for k in keys] + [(RF.COAL,),
                                         (cart["post"], cart["mid"], cart["ant"], RF.COAL, RF.FUSE, RF.FIB)]
    fr = RF.body_fractions(o["L"], sets, device=device).cpu().numpy()
    d = {}
    for q, key in enumerate(keys):
        fb = fr[q]
        m = (fb >= 0.25) & active
        mm = (fb > 0) & active
        d[key + "_p99"] = float(np.percentile(VM[m], 99)) if m.sum() >= 3 else 0.0
        w = fb[mm] * vol[mm]
        d[key + "_mean"] = float((VM[mm] * w).sum() / w.sum()) if mm.any() else 0.0
    Ft = o["res"]["F_on_talus"]
    fz = {k: max(Ft.get(k, [0, 0, 0])[2], 0.0) for k in ("ant", "mid", "post")}
    tot = sum(fz.values()) + 1e-9
    for k in fz:
        d["share_" + k] = float(fz[k] / tot)
    xyz = g.node_xyz[o["orig"]]
    bn = o["bone_nodes"]
    Un = np.nan_to_num(U)
    t, w, c = rigid_fit(xyz[bn["calcaneus"]], Un[bn["calcaneus"]])
    Urel = Un - (t[None, :] + np.cross(w[None, :], xyz - c[None, :]))
    tt, wt, _ = rigid_fit(xyz[bn["talus"]], Urel[bn["talus"]])
    d["talus_dz"] = float(tt[2])
    d["talus_dy"] = float(tt[1])
    d["talus_evert"] = float(np.rad2deg(wt[0]))
    d["nav_drop"] = float(-np.mean(Urel[bn["navicular"], 2]))
    d["heel_share"] = float(o["res"]["ground_Fz"]["heel"] / max(o["res"]["ground_total_Fz"], 1e-9))
    if dt is not None and dc is not None:
        kt = next(i for i, n in o["bnames"].items() if n == "talus")
        kc = next(i for i, n in o["bnames"].items() if n == "calcaneus")
        band = active & (np.asarray(dt) <= 6) & (np.asarray(dc) <= 6) & ((part == kt) | (part == kc) | (fr[6] > 0))
        d["tc_p99"] = float(np.percentile(VM[band], 99)) if band.sum() > 3 else 0.0
    cm = (fr[5] >= 0.25) & active
    d["coal_p99"] = float(np.percentile(VM[cm], 99)) if cm.sum() > 3 else 0.0
    d["coal_max"] = float(VM[cm].max()) if cm.sum() > 0 else 0.0
    return d, Urel
