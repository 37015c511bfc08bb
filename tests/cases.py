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
