"""The learned warm start's kernels on the GPU against exact references: the BF16 matrix-core
product against numpy on the same BF16-rounded operands (FP32 accumulation: relative 1e-5), the
whole network's forward and backward against the FP64 oracle (oracle/net_oracle.py, itself
pinned by central differences) on the same weights and inputs (BF16 activations: relative L2 a
few 1e-3 per layer, bounded at 3e-2), a short training run that lowers the loss, the cell <->
node maps, and the fallback rule of the warm-started solve."""
import time

import numpy as np
import pytest
import torch

from comech_amd import _lib
from comech_amd.learned import net as NN
from oracle import net_oracle as N

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a, dt):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=DEV)


@pytest.mark.parametrize("M,N,K,mode,ksplit", [(200, 16, 64, 0, 1), (1000, 64, 448, 1, 1), (333, 128, 3456, 0, 1),
                                               (16, 448, 4096, 2, 8), (128, 3456, 768, 2, 3), (77, 32, 864, 2, 1)])
def test_gemm_tn_matches_numpy(M, N, K, mode, ksplit):
    rng = np.random.default_rng(M + N)
    A = NN.to_bf16(rng.standard_normal((M, K)))
    B = NN.to_bf16(rng.standard_normal((N, K)))
    bias = rng.standard_normal(N).astype(np.float32)
    add = NN.to_bf16(rng.standard_normal((M, N)))
    ref = NN.from_bf16(A).astype(np.float64) @ NN.from_bf16(B).astype(np.float64).T
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    if mode == 2:
        Ad, Bd = _dev(A, torch.int16), _dev(B, torch.int16)     # held until the product has run
        C = _dev(np.full((M, N), 0.5, np.float32), torch.float32)
        _lib.check_nn(lib.cm_nn_gemm_tn(st, M, N, K, Ad.data_ptr(), K, Bd.data_ptr(), K, C.data_ptr(), N, None, None,
                                        0, 0, 2, ksplit, 1.0))
        got = C.cpu().numpy().astype(np.float64) - 0.5
    else:
        Ad, Bd, bd, addd = _dev(A, torch.int16), _dev(B, torch.int16), _dev(bias, torch.float32), _dev(add, torch.int16)
        C = torch.empty((M, N), dtype=torch.float32 if mode == 0 else torch.int16, device=DEV)
        _lib.check_nn(lib.cm_nn_gemm_tn(st, M, N, K, Ad.data_ptr(), K, Bd.data_ptr(), K, C.data_ptr(), N, bd.data_ptr(),
                                        addd.data_ptr(), N, 1, mode, 1, 1.0))
        ref = np.maximum(ref + bias + NN.from_bf16(add), 0.0)
        got = C.cpu().numpy() if mode == 0 else NN.from_bf16(C.cpu().numpy())
    tol = 1e-5 if mode != 1 else 8e-3                      # mode 1 rounds the result to BF16
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err <= tol, err


def _inputs(B, dims, seed=0):
    rng = np.random.default_rng(seed)
    f = NN.from_bf16(NN.to_bf16(rng.standard_normal((B,) + dims + (N.IN_CH,)))).astype(np.float64)
    t = rng.standard_normal((B,) + dims + (3,))
    mask = rng.uniform(size=(B,) + dims) < 0.8
    return f, t, mask


def test_network_forward_backward_match_oracle():
    """Same weights (BF16-representable) and features: the GPU network's output and every layer's
    weight and bias gradients against the FP64 oracle.  Biases are positive so every layer is
    live; a batch of 2 rasters checks that the convolutions do not mix rasters."""
    B, dims = 2, (16, 8, 8)
    rng = np.random.default_rng(5)
    P = N.init_params(seed=7)
    for name, cin, cout, _ in N.LAYERS:
        W, b = P[name]
        W[:] = NN.from_bf16(NN.to_bf16(W))
        b[:cout] = 0.1 + 0.05 * rng.standard_normal(cout)
    net = NN.Net(DEV, params=P)
    f, t, mask = _inputs(B, dims)
    out_o, C = N.forward(P, f)
    L_o, dout_o = N.loss(out_o, t, mask)
    G_o, _ = N.backward(P, C, dout_o)
    M = B * int(np.prod(dims))
    out = net.forward(_dev(NN.to_bf16(f.reshape(M, -1)), torch.int16), dims, B, keep=True)
    got = out.view(M, 16)[:, :3].cpu().numpy().reshape(out_o.shape)
    e_out = np.linalg.norm(got - out_o) / np.linalg.norm(out_o)
    part, dout, cnt = net.loss(out, _dev(t.reshape(M, 3).astype(np.float32), torch.float32),
                               _dev(mask.reshape(M).astype(np.uint8), torch.uint8), M)
    L = float(part.sum().item()) / (3 * cnt)
    net.zero_grad()
    net.backward(dout)
    G = net.grads()
    print(f"output rel L2 {e_out:.2e}, loss {L:.6f} vs {L_o:.6f}", flush=True)
    assert e_out < 3e-2 and abs(L - L_o) < 3e-2 * L_o
    # BF16 gradients pass through up to 15 layers (rounded at each) and ReLU masks flip where a
    # pre-activation is within rounding of zero: a few per cent on the first layers.  A wrong index,
    # tap or transposition gives cosines near 0.
    bad = []
    for name, cin, cout, _ in N.LAYERS:
        for q, what in ((0, "W"), (1, "b")):
            a, o = G[name][q], G_o[name][q]
            e = np.linalg.norm(a - o) / np.linalg.norm(o)
            cos = float((a * o).sum() / (np.linalg.norm(a) * np.linalg.norm(o)))
            print(f"{name}.{what}: rel L2 {e:.2e} cos {cos:.5f}", flush=True)
            if not (e < 1e-1 and cos > 0.995):
                bad.append((name, what, e, cos))
    assert not bad, bad


def test_training_lowers_the_loss():
    """200 Adam steps on 4 small rasters (fit fixed random-but-smooth targets) lower the masked MSE
    at least 5x: the forward, backward and optimizer kernels train the network."""
    B, dims = 4, (16, 16, 8)
    rng = np.random.default_rng(9)
    f, _, mask = _inputs(B, dims, seed=9)
    g = np.stack(np.meshgrid(*[np.linspace(0, 1, n) for n in dims], indexing="ij"), -1)
    t = np.stack([np.stack([np.sin(3 * g[..., 0] + b) * g[..., 2], np.cos(2 * g[..., 1] - b), g[..., 0] * g[..., 1]],
                           -1) for b in range(B)])
    M = B * int(np.prod(dims))
    net = NN.Net(DEV, seed=2)
    Fd = _dev(NN.to_bf16(f.reshape(M, -1)), torch.int16)
    Td = _dev(t.reshape(M, 3).astype(np.float32), torch.float32)
    Md = _dev(mask.reshape(M).astype(np.uint8), torch.uint8)
    losses = []
    for it in range(200):
        out = net.forward(Fd, dims, B, keep=True)
        part, dout, cnt = net.loss(out, Td, Md, M)
        net.zero_grad()
        net.backward(dout)
        net.adam(2e-3)
        if it % 20 == 0 or it == 199:
            losses.append(float(part.sum().item()) / (3 * cnt))
    print("loss", [round(x, 5) for x in losses], flush=True)
    assert losses[-1] < losses[0] / 5


def test_cell_node_maps_round_trip():
    """A displacement field linear in the node coordinates: the cell means (targets) of the exact
    field, mapped back to nodes, reproduce the field at every node inside the active region
    (split copies included), and the map scales with the load."""
    from comech_amd import synth
    from comech_amd.learned import data as D
    c = synth.hindfoot_raster(dims=(24, 16, 16), h=3.0)
    desc = synth.level0_description(c)
    enc = D.encode(desc, DEV)
    xyz = c["grid"].node_xyz[c["orig"]]
    u = np.stack([0.01 * xyz[:, 0], -0.02 * xyz[:, 1] + 0.003 * xyz[:, 2], np.full(len(xyz), 0.01)], 1)
    x = _dev(u.ravel(), torch.float64)
    t = D.targets(enc, x, 1.0).view(-1, 3)
    pred = torch.zeros((t.shape[0], 16), dtype=torch.float32, device=DEV)
    pred[:, :3] = t
    back = D.to_nodes(enc, pred, 1.0).view(-1, 3).cpu().numpy()
    # a node's mean over its cells of the cell means equals the node value only for fields constant
    # in each cell's neighbourhood; a linear field leaves the interior exact
    used = c["used"]
    cnt = np.bincount(c["cn"][c["active"]].ravel(), minlength=len(used))
    sel = (cnt == 8) & used                      # nodes inside the active region (8 active cells)
    assert sel.sum() > 100
    assert np.abs(back[sel] - u[sel]).max() < 1e-6 * np.abs(u).max()
    assert not back[~used].any()
    F = enc["ftot"]
    assert F > 0 and abs(F - 700.0) < 1e-6


def test_fallback_rule_hands_over_to_plain():
    """A bad initial guess with an unreachable drop (1e12 within 5 iterations) stops at iteration 5
    with flag -2; the caller's plain solve converges; a reachable rule (drop 1) never triggers."""
    from comech_amd import synth
    from comech_amd.batch import build_system, rhs
    c = synth.hindfoot_raster(dims=(24, 16, 16), h=3.0)
    desc = synth.level0_description(c)
    s = build_system(desc, device=DEV)
    b = rhs(desc, DEV)
    x0 = torch.randn_like(b) * 1e-2
    x0[~torch.as_tensor(np.repeat(s.used_mask(0), 3), device=DEV)] = 0
    s.set_fallback(5, 1e12)
    x, info = s.solve(b, x0, tol=1e-8, maxiter=2000, mode="ref")
    assert info["flag"] == -2 and info["iters"] == 5
    s.set_fallback(0)
    x, info = s.solve(b, None, tol=1e-8, maxiter=2000, mode="ref")
    assert info["flag"] == 0
    s.set_fallback(5, 1.0)
    x, info = s.solve(b, x0, tol=1e-8, maxiter=2000, mode="ref")
    assert info["flag"] == 0
    s.close()


def test_trained_warm_start_on_a_held_out_foot():
    """The committed network on a foot of its held-out split (96x64x64, 8 actions): the prediction
    is within 5 % (relative L2) of the solution; projected onto its coarse space it removes at least
    15 % of the energy-norm error of a zero start (measured: 25 %) where the bones' rigid modes
    alone remove none; the learned solve converges to the plain answer without a fallback, in at
    most one residual check (5 iterations) more per system on average than the plain solve.
    Iterations and wall times are printed: at rtol 1e-8 the iteration counts are equal or within a
    check of each other (the removed error is the smooth part the V-cycle removes cheaply anyway),
    and the prediction plus projection cost time, so neither an iteration nor a wall-time gain is
    claimed."""
    from comech_amd.batch import build_system, rhs
    from comech_amd.learned import load_default
    from comech_amd.learned import train as TR
    from comech_amd.learned.data import encode
    from comech_amd.learned.warmstart import CoarseSpace, learned_solve, predict
    seed = TR.split_feet(500)[1][0]
    descs = TR.foot_inputs(seed)
    net = load_default(DEV)
    tot_p = tot_l = 0
    for d in descs:
        s = build_system(d, device=DEV)
        b = rhs(d, DEV)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xp, ip = s.solve(b, tol=1e-8, maxiter=3000, mode="ref")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        enc = encode(d, DEV)
        xl, il = learned_solve(s, d, b, net, tol=1e-8, mode="ref", enc=enc)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        pred = predict(net, enc)
        rel_pred = float(torch.linalg.vector_norm(pred - xp) / torch.linalg.vector_norm(xp))
        anorm = lambda e: float(torch.sqrt(torch.dot(e, s.apply(e))))
        g = CoarseSpace(enc, pred).galerkin(s, b)
        a_err = anorm(xp - g) / anorm(xp)
        g_rigid = CoarseSpace(enc, torch.zeros_like(pred)).galerkin(s, b)
        a_rigid = anorm(xp - g_rigid) / anorm(xp)
        rel = float(torch.linalg.vector_norm(xl - xp) / torch.linalg.vector_norm(xp))
        print(f"{d['action']}: iterations plain {ip['iters']} learned {il['iters']}, ms plain "
              f"{(t1 - t0) * 1e3:.1f} learned {(t2 - t1) * 1e3:.1f} (guess {il['t_guess_ms']:.1f}), prediction "
              f"rel {rel_pred:.4f}, energy error {a_err:.3f} (rigid modes only {a_rigid:.3f})", flush=True)
        s.close()
        assert rel_pred < 0.05
        assert a_err < 0.85 and a_rigid > 0.95
        assert il["flag"] == 0 and not il["fallback"] and rel < 1e-6
        tot_p += ip["iters"]
        tot_l += il["iters"]
    print(f"total iterations: plain {tot_p}, learned {tot_l}", flush=True)
    assert tot_l <= tot_p + 5 * len(descs)
