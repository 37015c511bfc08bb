"""Drop-in for the reference's engine interface ``fe/simulate.py``.

``simulate(foot, act, ...)`` keeps the reference signature (fe/simulate.py:15-20) but needs the
atlas template ``rb_template.npz`` and its geometry pipeline, which are not part of this solve
path; it raises ``NotImplementedError``.  ``simulate_raster`` is the sibling entry for voxel
rasters with springs (the synthetic configurations): it runs ``solve_cn``-style split-node FE on
the GPU and returns the reference's result keys for the FE part (``res``, ``iters``, ``ndof``,
``U``, ``VM``, ``SIG``, ``active``, ``timing``).
"""
import time

import numpy as np

from ..rb import splitfe as SF

BODY_LOAD = 360.0


def simulate(foot, act, fine=1.0, mat_override=None, return_fields=True, method="gmg", tol=1e-6, frac_hook=None,
             U0=None):
    raise NotImplementedError("simulate(foot, act) needs the atlas template rb_template.npz and its geometry "
                              "pipeline; use simulate_raster for voxel rasters")


def simulate_raster(case, tol=1e-8, return_fields=True, **opts):
    """case: dict with grid, lam, mu, active, cn, orig, fixed, loads, extra (COO over node DOFs or
    the reference callback), as produced by comech_amd.synth."""
    t0 = time.time()
    g = case["grid"]
    U, info, R = SF.solve_cn(g, case["lam"], case["mu"], case["active"], case["cn"], case["orig"], case["fixed"],
                             case["loads"], extra=case.get("extra_cb"), tol=tol, **opts)
    t1 = time.time()
    act = np.flatnonzero(case["active"])
    sig, vm, _ = SF.cell_stress_cn(g, case["lam"], case["mu"], U, act, case["cn"])
    VM = np.zeros(g.ncell, np.float32); VM[act] = vm
    SIG = np.zeros((g.ncell, 6), np.float32); SIG[act] = sig
    res = {"ndof": int(info["ndof"]), "iters": info["iters"], "flag": info["flag"], "relres": info["relres"],
           "t_solve_s": round(t1 - t0, 3), "levels": info["levels"]}
    out = dict(res=res, iters=info["iters"], ndof=info["ndof"], timing=(0.0, round(t1 - t0, 3)))
    if return_fields:
        out.update(U=U, VM=VM, SIG=SIG, active=np.asarray(case["active"], bool), R=R)
    return out
