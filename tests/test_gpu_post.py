"""GPU post-processing and drop-in interoperability: sample_fe and sample_nodes (rb/rbsim.py:110-161) against the
oracle, and solve_cn called with a reference-shaped Grid object (the oracle's restatement of
fe/solver.Grid, which has no helper methods of the product's Grid)."""
import numpy as np
import pytest

from comech_amd.rb import rbsim as RS
from comech_amd.rb import splitfe as SF
from oracle import fe_oracle as O
from tests import cases as C

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_sample_fe_matches_oracle():
    c = C.hindfoot((24, 16, 16), 3.0)
    og = C.oracle_grid(c)
    rng = np.random.default_rng(11)
    nn = len(c["orig"])
    Urel = rng.standard_normal((nn, 3))
    Urel[rng.uniform(size=nn) < 0.05] = np.nan                 # NaN entries count as 0
    VM = rng.uniform(0, 5, og.ncell).astype(np.float32)
    SIG = rng.standard_normal((og.ncell, 6)).astype(np.float32)
    # analysis grid that pokes outside the FE box on two sides
    gm = O.Grid(np.arange(-5.0, 75.0, 8.0), np.arange(2.0, 50.0, 8.0), np.arange(-3.0, 52.0, 8.0))
    vm_o, u_o, cov_o, sig_o = O.sample_fe(og, c["cn"], c["active"], VM, SIG, Urel, gm)
    o = dict(g=og, cn=c["cn"], active=c["active"], VM=VM, SIG=SIG)
    vm_g, u_g, cov_g = RS.sample_fe(o, gm, Urel, device=DEV)
    assert np.array_equal(cov_g, cov_o)
    assert np.abs(vm_g - vm_o).max() <= 1e-5 * np.abs(vm_o).max()
    assert np.abs(u_g - u_o).max() <= 1e-5 * np.abs(u_o).max()
    assert np.abs(RS.sample_fe.sig - sig_o).max() <= 1e-5 * np.abs(sig_o).max()


@pytest.mark.parametrize("sparse", [False, True])
def test_sample_nodes_matches_oracle(sparse):
    """rb/rbsim.py:143-161 on the GPU: bit-exact float32 against the oracle (scipy feature transform),
    split copies in the connectivity, NaN (unused) entries, an analysis grid reaching outside the FE
    box; ``sparse`` keeps 3 % of the active cells, so most nodes resolve a nearest cell with ties."""
    c = C.hindfoot((24, 16, 16), 3.0)
    og = C.oracle_grid(c)
    rng = np.random.default_rng(12)
    act = c["active"].copy()
    if sparse:
        act &= rng.uniform(size=act.size) < 0.03
    nn = len(c["orig"])
    Urel = rng.standard_normal((nn, 3))
    Urel[rng.uniform(size=nn) < 0.05] = np.nan
    gm = O.Grid(np.arange(-7.0, 80.0, 6.0), np.arange(-2.0, 53.0, 5.0), np.arange(1.0, 56.0, 7.0))
    u_o = O.sample_nodes(og, c["cn"], act, Urel, gm)
    u_g = RS.sample_nodes(dict(g=og, cn=c["cn"], active=act), gm, Urel, device=DEV)
    assert u_g.shape == u_o.shape and u_g.dtype == np.float32
    assert np.array_equal(u_g, u_o)


def test_sample_nodes_rejects_empty_mask():
    c = C.hindfoot((24, 16, 16), 3.0)
    og = C.oracle_grid(c)
    gm = O.Grid(np.arange(0.0, 70.0, 10.0), np.arange(0.0, 50.0, 10.0), np.arange(0.0, 50.0, 10.0))
    with pytest.raises(ValueError):
        RS.sample_nodes(dict(g=og, cn=c["cn"], active=np.zeros(og.ncell, bool)), gm,
                        np.zeros((len(c["orig"]), 3)), device=DEV)


def test_solve_cn_with_reference_shaped_grid():
    c = C.hindfoot((24, 16, 16), 3.0)
    og = C.oracle_grid(c)
    U, info, R = SF.solve_cn(og, c["lam"], c["mu"], c["active"], c["cn"], c["orig"], c["fixed"], c["loads"],
                             extra=c["extra_cb"], tol=1e-10, device=DEV)
    Uo, _, _ = O.solve_cn(og, c["lam"], c["mu"], c["active"], c["cn"], c["orig"], c["fixed"], c["loads"],
                          extra=c["extra_cb"], tol=1e-12, maxiter=2000)
    err = np.linalg.norm(np.nan_to_num(U - Uo)) / np.linalg.norm(np.nan_to_num(Uo))
    assert info["flag"] == 0 and err <= 1e-6


def test_bank_pack_and_files(tmp_path):
    """Bank arrays packed on the GPU equal the reference's numpy expressions (exp/bank.py:116-118:
    vm and the cell-mean displacement as float16, lab as int8, round(255 frac) as uint8), and the
    written foot files load back with allow_pickle=False."""
    import json
    from comech_amd import bank
    from comech_amd.grid import Grid
    rng = np.random.default_rng(5)
    # 120,000 cells: the float32 summation order of the cell mean differs from the index order in
    # about half of the cells, so a wrong order fails here (float16 of the mean differs in ~1e-4)
    g = Grid(np.arange(61.0), np.arange(51.0), np.arange(41.0))
    Urel = rng.standard_normal((g.nnode, 3)).astype(np.float32) * 3
    vm = rng.uniform(0, 40, g.ncell).astype(np.float32)
    lab = rng.integers(-1, 9, g.ncell)
    frac = rng.dirichlet(np.ones(10), g.ncell).astype(np.float32)
    frac[0, :2] = [0.5 / 255, 1.5 / 255]                  # rounding ties
    p = bank.pack_action(vm, Urel, g.cell_nodes, lab, frac, device=DEV)
    assert np.array_equal(p["vm"], vm.astype(np.float16))
    v = Urel[g.cell_nodes]
    uc = v.mean(1).astype(np.float16)
    pairwise = ((((v[:, 0] + v[:, 1]) + (v[:, 2] + v[:, 3])) + ((v[:, 4] + v[:, 5]) + (v[:, 6] + v[:, 7])))
                / np.float32(8)).astype(np.float16)
    assert not np.array_equal(pairwise, uc)              # the sample separates the two orders
    assert np.array_equal(p["ucell"], uc)
    assert np.array_equal(p["lab"], lab.astype(np.int8))
    assert np.array_equal(p["frac"], np.round(frac * 255).astype(np.uint8))
    rec = [dict(k=k, meta="baseline", act={"load": 1.0}, eff={"heel_share": 0.5}, iters=80, time=0.1, fe={})
           for k in range(2)]
    npz, js = bank.write_foot(str(tmp_path), "train", 7, {"ctype": 1}, rec, [p, p])
    z = np.load(npz)
    assert z["vm"].shape == (2, g.ncell) and z["vm"].dtype == np.float16 and z["ucell"].shape == (2, g.ncell, 3)
    assert z["lab"].dtype == np.int8 and z["frac"].dtype == np.uint8
    m = json.load(open(js))
    assert m["fid"] == 7 and m["split"] == "train" and len(m["actions"]) == 2
