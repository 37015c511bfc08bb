"""Shared test cases: product inputs and the oracle's assembled systems for the same seeds."""
import functools

import numpy as np

from comech_amd import synth
from oracle import fe_oracle as O


@functools.lru_cache(maxsize=None)
def cube(n=10):
    return synth.cube_inclusions(n=n, h=1.0)


@functools.lru_cache(maxsize=None)
def hindfoot(dims=(40, 24, 24), h=2.0):
    return synth.hindfoot_raster(dims=dims, h=h)


def oracle_grid(case):
    g = case["grid"]
    return O.Grid(g.xe, g.ye, g.ze)


@functools.lru_cache(maxsize=None)
def oracle_system1(n=10):
    c = cube(n)
    _, info = O.solve(oracle_grid(c), c["lam"], c["mu"], c["active"], c["springs"], c["loads"], method="none")
    return info


@functools.lru_cache(maxsize=None)
def oracle_system2(dims=(40, 24, 24), h=2.0):
    c = hindfoot(dims, h)
    _, info, _ = O.solve_cn(oracle_grid(c), c["lam"], c["mu"], c["active"], c["cn"], c["orig"], c["fixed"],
                            c["loads"], extra=c["extra_cb"], method="none")
    return info


def random_full(nn, used, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((nn, 3))
    u[~used] = 0.0
    return u.ravel()


def make_case(kind, kw):
    """Product inputs of one parity case: 'cube' (cube_inclusions), 'hindfoot' (hindfoot_raster)
    or 'batch' (batch_case)."""
    if kind == "cube":
        return synth.cube_inclusions(**kw)
    if kind == "batch":
        return synth.batch_case(**kw)
    return synth.hindfoot_raster(**kw)


def oracle_answer(spec):
    """FP64 oracle of one case (run in a worker process): the reference CPU solver (assembled CSR,
    Galerkin GMG-PCG, SuperLU coarse solve) to rtol 1e-12, von Mises of the active cells and, for
    split-node cases, the ground reactions; also the assembled system for residual checks."""
    kind, kw, mg_kw = spec
    c = make_case(kind, kw)
    og = oracle_grid(c)
    act = np.flatnonzero(c["active"])
    if kind == "cube":
        U, info = O.solve(og, c["lam"], c["mu"], c["active"], c["springs"], c["loads"], method="gmg", tol=1e-12,
                          **mg_kw)
        R = None
        _, vm, _ = O.cell_stress(og, c["lam"], c["mu"], U, act)
    else:
        U, info, R = O.solve_cn(og, c["lam"], c["mu"], c["active"], c["cn"], c["orig"], c["fixed"], c["loads"],
                                extra=c["extra_cb"], tol=1e-12, maxiter=4000, **mg_kw)
        _, vm, _ = O.cell_stress(og, c["lam"], c["mu"], U, act, cn=c["cn"])
    return dict(U=U, vm=vm, R=R, K=info["K"].tocsr(), f=info["f"], used=np.asarray(info["used"]),
                iters=info.get("iters"))


def oracle_answers(specs, workers=8):
    """oracle_answer over many cases on a pool of CPU worker processes (spawned; they import
    numpy/scipy and never touch the GPU)."""
    import multiprocessing as mp
    import os
    n = max(1, min(workers, len(specs), os.cpu_count() or 1))
    if n == 1:
        return [oracle_answer(s) for s in specs]
    with mp.get_context("spawn").Pool(n) as pool:
        return pool.map(oracle_answer, specs, chunksize=1)
