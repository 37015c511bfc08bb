"""The allocation-free entry point cm_sample_nodes_noalloc (work tables in the library's
per-device scratch) against cm_sample_nodes (work arrays passed in): bit-equal float32 results on
two calls in a row, the second on a smaller input, so stale scratch contents from the first call
would show; a size above the scratch limit is rejected with an error, not run."""
import ctypes

import numpy as np
import pytest
import torch

from comech_amd import _lib
from comech_amd.rb import rbsim as RS
from oracle import fe_oracle as O
from tests import cases as C

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _case(dims, h, seed, sparse):
    c = C.hindfoot(dims, h)
    og = C.oracle_grid(c)
    rng = np.random.default_rng(seed)
    act = c["active"].copy()
    if sparse:
        act &= rng.uniform(size=act.size) < 0.05
    nn = len(c["orig"])
    Urel = rng.standard_normal((nn, 3))
    Urel[rng.uniform(size=nn) < 0.05] = np.nan
    L = [e[-1] for e in (og.xe, og.ye, og.ze)]
    gm = O.Grid(np.linspace(-4.0, L[0] + 3.0, 13), np.linspace(-2.0, L[1] + 1.0, 9), np.linspace(1.0, L[2] + 2.0, 8))
    return dict(g=og, cn=c["cn"], active=act), gm, Urel


def test_noalloc_equals_workspace_entry_point_twice():
    first = _case((40, 24, 24), 2.0, 21, sparse=True)
    second = _case((24, 16, 16), 3.0, 22, sparse=False)     # smaller FE grid and fewer query layers
    for o, gm, U in (first, second):
        ref = RS.sample_nodes(o, gm, U, device=DEV)
        got = RS.sample_nodes(o, gm, U, device=DEV, noalloc=True)
        assert got.shape == ref.shape and got.dtype == np.float32
        assert np.array_equal(got, ref)
        # and both equal the oracle (scipy feature transform)
        assert np.array_equal(got, O.sample_nodes(o["g"], o["cn"], o["active"], U, gm))


def test_noalloc_rejects_sizes_above_the_scratch_limit():
    """nq * nx * ny above CM_SAMPLE_NODES_MAX_COLUMNS fails on the host check, before any launch
    (the pointers are never dereferenced)."""
    lib = _lib.load()
    torch.cuda.init()
    nx = ny = 2048
    nq = 2                                            # 2 * 2048 * 2048 = 8,388,608 > 4,194,304
    null = ctypes.c_void_p(0)
    rc = lib.cm_sample_nodes_noalloc(null, 4, 4, 4, *[null] * 6, nq, null, null, nx, ny, 8, *[null] * 6, null)
    assert rc == -1
    assert b"exceeds the scratch limit" in lib.cm_last_error()
