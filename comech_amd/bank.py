"""Bank output writer (exp/bank.py:106-131): the engine's results for one foot under its K
actions as ``foot_XXXX.npz`` (vm float16 (K, ncell), ucell float16 (K, ncell, 3), lab int8
(K, ncell), frac uint8 (K, ncell, ncls)) and ``foot_XXXX.json`` (fid, split, foot, one record per
action), the files gpu/cache.py:40-67 reads.  The per-action arrays are packed on the GPU
(``cm_bank_pack``: FP16 conversion, the cell mean of the relative displacement over the analysis
cell's 8 nodes, exp/bank.py:116, and round(255 frac)); writing the files is host IO.
"""
import ctypes
import json
import os

import numpy as np
import torch

from ._lib import check, load
from .system import require_gpu


def pack_action(vm, Urel, cell_nodes, lab, frac, device=None):
    """One action's bank arrays: vm (ncell,) float, Urel (nnode, 3) relative nodal displacement of
    the analysis grid, cell_nodes (ncell, 8), lab (ncell,), frac (ncell, ncls).  Returns numpy
    arrays (vm float16, ucell float16 (ncell, 3), lab int8, frac uint8)."""
    dev = require_gpu(device)
    lib = load()
    t = lambda a, dt: (a.to(device=dev, dtype=dt).contiguous() if isinstance(a, torch.Tensor)
                       else torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev))
    vmd = t(vm, torch.float32)
    ud = t(np.nan_to_num(np.asarray(Urel, np.float32)) if not isinstance(Urel, torch.Tensor) else Urel, torch.float32)
    cnd = t(cell_nodes, torch.int64)
    labd = t(lab, torch.int32)
    frd = t(frac, torch.float32)
    ncell = int(vmd.shape[0])
    ncls = int(frd.shape[1]) if frd.dim() == 2 else 0
    if cnd.shape != (ncell, 8) or labd.shape[0] != ncell or (ncls and frd.shape[0] != ncell):
        raise ValueError("vm, cell_nodes, lab and frac must describe the same cells")
    if ncell and (int(cnd.max().item()) >= ud.shape[0] or int(cnd.min().item()) < 0):
        raise ValueError("cell_nodes index outside Urel")
    vm16 = torch.empty(ncell, dtype=torch.float16, device=dev)
    uc16 = torch.empty((ncell, 3), dtype=torch.float16, device=dev)
    lab8 = torch.empty(ncell, dtype=torch.int8, device=dev)
    fr8 = torch.empty((ncell, ncls), dtype=torch.uint8, device=dev)
    vp = lambda x: ctypes.c_void_p(x.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    check(lib.cm_bank_pack(st, ncell, ncls, vp(vmd), vp(ud), vp(cnd), vp(labd), vp(frd), vp(vm16), vp(uc16), vp(lab8),
                           vp(fr8)))
    return dict(vm=vm16.cpu().numpy(), ucell=uc16.cpu().numpy(), lab=lab8.cpu().numpy(), frac=fr8.cpu().numpy())


def write_foot(root, split, fid, foot, records, packs):
    """Write ``root/split/foot_{fid:04d}.npz`` and ``.json`` (exp/bank.py:123-128).  ``records``: one
    dict per action (k, meta, act, eff, iters, time, fe); ``packs``: pack_action results in action
    order.  Returns the two paths."""
    d = os.path.join(root, split)
    os.makedirs(d, exist_ok=True)
    base = os.path.join(d, f"foot_{int(fid):04d}")
    np.savez_compressed(base + ".npz", vm=np.stack([p["vm"] for p in packs]),
                        ucell=np.stack([p["ucell"] for p in packs]), lab=np.stack([p["lab"] for p in packs]),
                        frac=np.stack([p["frac"] for p in packs]))
    with open(base + ".json", "w") as fh:
        json.dump(dict(fid=int(fid), split=split, foot=foot, actions=list(records)), fh)
    return base + ".npz", base + ".json"


__all__ = ["pack_action", "write_foot"]
