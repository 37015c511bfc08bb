"""Drop-in for ``sample_fe`` of the reference's ``rb/rbsim.py:110-140`` on the GPU.

Every analysis cell averages von Mises stress, stress and displacement over sub^3 sample points in
the FE cells; samples outside the FE box or in inactive FE cells are excluded.  The host builds
the per-axis sample tables (cell index, local coordinate, inside flag); the averaging runs in the
``k_sample_fe`` kernel.  Returns (vm, u, cover) and sets ``sample_fe.sig`` like the reference.
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from ..system import require_gpu


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
    a = t(np.asarray(act, np.uint8), torch.uint8)
    c = t(np.asarray(cn, np.int64).reshape(-1), torch.int64)
    vm_in = t(np.asarray(o["VM"], np.float32), torch.float32)
    sg_in = t(np.asarray(o["SIG"], np.float32).reshape(-1), torch.float32)
    U = t(np.asarray(Urel, float).reshape(-1), torch.float64)
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
