"""Drop-in for the engine preparation of ``rb/rbfe.py`` on the GPU: lattice material mixing
(``materials``, rb/rbfe.py:75-134), bone shares (``bone_share``, rb/replica3.py:17-29) and the
bone part of every cell (rb/rbfe.py:140-141).

A lattice ``V`` (as built by the reference's voxeliser) holds ``L`` -- (ncell, ns^3) body labels
of the sub-samples of every cell in reference cell order (i ny + j) nz + k -- and ``cls`` -- the
tissue class of every body label.  ``materials`` returns the same (E, nu, active, F, ser) as the
reference; the per-cell work (class counts, the mixing rule, 3x3x3 joint detection) runs in
``cm_lattice_materials``.  There is no CPU path.

The class list, moduli, bone and cartilage body tables are the engine's definitions
(rb/rbgeom.py:35-49, rb/rbfe.py:29) and are kept identical.
"""
import ctypes

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
# rb/rbfe.py:29
ROLES = {"bones": dict(BONES), "cart": dict(CART), "bonded_extra": [COAL, FUSE]}


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


__all__ = ["COAL", "FIB", "INTERP", "FUSE", "GAP", "BONES", "CART", "CLASSES", "MAT", "ROLES", "bone_groups",
           "lattice_prep", "materials", "bone_share"]
