"""GPU engine preparation against the oracle: lattice material mixing, bone shares and part
(rb/rbfe.py:75-141, rb/replica3.py:17-29) and split connectivity (rb/splitfe.py:261-305).
Booleans, fractions, shares, parts and node numbering bit-exact; (E, nu) to 1e-12 relative
(FP64, the same formulas evaluated in the same order)."""
import numpy as np
import pytest

from comech_amd.grid import Grid
from comech_amd.rb import rbfe as RF
from comech_amd.rb import splitfe as SF
from oracle import fe_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

# body labels -> classes: bones (cortical 1 / cancellous 2), cartilage bodies (3), fascia (4), soft
# (5), plate (6), coalition bars (cortical, bonded to the bones they touch), fibrous (7), fat (8),
# gap (9), and one unnamed bone body (its own group)
CLS = {1: 1, 53: 2, 2: 1, 52: 2, 3: 1, 54: 2, 40: 1, 45: 2, 5: 3, 8: 3, 12: 3, 60: 4, 61: 5, 62: 6,
       RF.COAL: 1, RF.FIB: 7, RF.INTERP: 8, RF.FUSE: 1, RF.GAP: 9, 77: 2}
LABELS = [0, 0, 1, 53, 2, 52, 3, 54, 40, 45, 5, 8, 12, 60, 61, 61, 62, RF.COAL, RF.FIB, RF.INTERP, RF.FUSE, RF.GAP, 77]


def _lattice(dims, ns3, seed, blobs=True):
    rng = np.random.default_rng(seed)
    nc = int(np.prod(dims))
    L = rng.choice(LABELS, size=(nc, ns3))
    if blobs:   # spatially coherent bone regions so that joints and parts occur
        nx, ny, nz = dims
        i, j, k = np.unravel_index(np.arange(nc), dims)
        L[(i < nx // 2) & (rng.uniform(size=nc) < 0.8)] = 1
        L[(i >= nx // 2) & (j < ny // 2) & (rng.uniform(size=nc) < 0.8)] = 52
    L[rng.uniform(size=nc) < 0.1] = 0            # empty cells
    return dict(L=L, cls=CLS, xe=np.arange(dims[0] + 1.0), ye=np.arange(dims[1] + 1.0), ze=np.arange(dims[2] + 1.0))


@pytest.mark.parametrize("seed,ns3,blobs", [(0, 27, True), (1, 27, False), (2, 8, True), (3, 8, False)])
def test_lattice_materials_match_oracle(seed, ns3, blobs):
    dims = (9, 7, 6)
    V = _lattice(dims, ns3, seed, blobs)
    o = RF.lattice_prep(V, cart_E=12.0, nu_soft=0.48, device=DEV)
    E, nu, act, F, ser = O.lattice_materials(V["L"], CLS, RF.ROLES, dims, cart_E=12.0, nu_soft=0.48)
    S, names = O.bone_share(V["L"], CLS, RF.ROLES)
    part = O.lattice_part(S, F)
    assert np.array_equal(o["active"], act) and np.array_equal(o["ser"], ser)
    assert np.array_equal(o["F"], F) and np.array_equal(o["S"], S) and o["names"] == names
    assert np.array_equal(o["part"], part)
    assert np.allclose(o["E"], E, rtol=1e-12, atol=0) and np.allclose(o["nu"], nu, rtol=1e-12, atol=0)
    assert act.any() and ser.any() and (~act).any()
    assert (part > 0).any() or not blobs          # random labels alone rarely fill half a cell with one bone


@pytest.mark.parametrize("seed", range(4))
def test_split_connectivity_gpu_bit_exact(seed):
    """Random parts, activity and shares (ties and zero-share cells included)."""
    rng = np.random.default_rng(200 + seed)
    shape = (9, 7, 8)
    g, og = Grid(*(np.arange(n + 1.0) for n in shape)), O.Grid(*(np.arange(n + 1.0) for n in shape))
    nb = 5
    part = rng.integers(0, nb, g.ncell)
    active = rng.uniform(size=g.ncell) > 0.1
    share = rng.uniform(size=(g.ncell, nb)).astype(np.float32)
    share[rng.uniform(size=g.ncell) < 0.3] = 0
    share[:, 0] = 0
    share[rng.uniform(size=g.ncell) < 0.2, 2] = share[rng.uniform(size=g.ncell) < 1.0, 1][0]   # ties
    cn, orig, ns = SF.split_connectivity_gpu(g, active, part, share, device=DEV)
    cno, origo, nso = O.split_connectivity(og, active, part, share)
    assert ns == nso and np.array_equal(orig, origo) and np.array_equal(cn, cno)


def test_split_connectivity_gpu_hindfoot_raster():
    """The synthetic hindfoot raster's bones: GPU == the product's host form == oracle."""
    from comech_amd import synth
    c = synth.config2(dims=(40, 24, 24), h=2.0)
    g = c["grid"]
    cn, orig, ns = SF.split_connectivity_gpu(g, c["active"], c["part"], c["bone_share"], device=DEV)
    cnh, origh, nsh = SF.split_connectivity_fast(g, c["active"], c["part"], c["bone_share"])
    assert ns == nsh > 0 and np.array_equal(orig, origh) and np.array_equal(cn, cnh)


def test_lattice_then_split_chain():
    """Lattice prep (shares, part, activity) feeding the GPU split connectivity, as rb/rbfe.py:139-142
    chains them, equals the oracle chain."""
    dims = (10, 8, 7)
    V = _lattice(dims, 27, 9)
    o = RF.lattice_prep(V, device=DEV)
    g, og = Grid(*(np.arange(n + 1.0) for n in dims)), O.Grid(*(np.arange(n + 1.0) for n in dims))
    cn, orig, ns = SF.split_connectivity_gpu(g, o["active"], o["part"], o["S"], device=DEV)
    E, nu, act, F, ser = O.lattice_materials(V["L"], CLS, RF.ROLES, dims)
    S, _ = O.bone_share(V["L"], CLS, RF.ROLES)
    cno, origo, nso = O.split_connectivity(og, act, O.lattice_part(S, F), S)
    assert ns == nso and np.array_equal(orig, origo) and np.array_equal(cn, cno)
