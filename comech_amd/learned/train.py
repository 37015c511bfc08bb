"""Training of the learned warm start on solves the engine itself generates.

    python -m comech_amd.learned.train --feet 500 --epochs 12 --out weights.npz --summary train.json

Data: the hindfoot family (the anatomy of the batched set) at 96x64x64 voxels of 0.5 mm -- seeded
feet (seed 4000 + i, osteotomy angle seed + 500) under the 8 surgical actions, each solved on the
GPU by the project's GMG-PCG (mode ref, rtol 1e-8).  The split is 90/10 by foot (a seeded
permutation), so no test foot's anatomy is seen in training.  Input: the 16 raster channels of
comech_amd.learned.data; target: the cell-mean displacement divided by the total load, times a
fixed scale (1 / its RMS over the training set's active cells).  Loss: masked mean-squared error.
Optimiser: Adam on FP32 master weights with BF16 compute copies, a linear warm-up and a cosine
decay.  Every FLOP of the forward and backward passes and of the optimiser runs in the project's
kernels (include/comech_learned.h).  At the end the held-out feet are solved with and without the
warm start, and the weights and a JSON summary are written.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

DIMS = (96, 64, 64)
H = 0.5
SEED0 = 4000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def foot_inputs(seed, dims=DIMS, h=H):
    """CPU worker: level-0 descriptions of one foot under the 8 actions."""
    from comech_amd import synth
    fam = synth.hindfoot_family(dims, h, seed, angle_seed=seed + 500)
    return [synth.level0_description(fam[a]) for a in synth.ACTIONS]


def split_feet(n_feet, seed=SEED0, test_frac=0.1):
    """Seeded 90/10 split of the foot seeds (train seeds, test seeds)."""
    perm = np.random.default_rng(seed).permutation(n_feet)
    nt = max(1, int(round(test_frac * n_feet)))
    return sorted(SEED0 + perm[nt:]), sorted(SEED0 + perm[:nt])


def gather_inputs(seeds, workers):
    """Descriptions of every (foot, action), in seed order, from a pool of host processes."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        per_foot = list(ex.map(foot_inputs, seeds, chunksize=1))
    return [d for foot in per_foot for d in foot]


def solve_set(descs, dev, conc=16, tol=1e-8):
    """Solve every description on the GPU; returns device tensors F [N][V*16] (BF16 bits), T [N][V*3]
    (cell-mean displacement / total load), M [N][V] (active), and the solver's iterations."""
    import torch
    from comech_amd.batch import build_system, rhs, run_batch
    from comech_amd.learned.data import encode, targets
    V = int(np.prod(DIMS))
    N = len(descs)
    F = torch.empty((N, V * 16), dtype=torch.int16, device=dev)
    T = torch.empty((N, V * 3), dtype=torch.float32, device=dev)
    M = torch.empty((N, V), dtype=torch.uint8, device=dev)
    iters = [0] * N

    def job(i):
        def run():
            s = build_system(descs[i], device=dev)
            b = rhs(descs[i], dev)
            x, info = s.solve(b, tol=tol, maxiter=3000, mode="ref")
            s.close()
            if info["flag"]:
                raise RuntimeError(f"training solve {i} did not converge: {info}")
            enc = encode(descs[i], dev)
            F[i].copy_(enc["F"])
            T[i].copy_(targets(enc, x, 1.0))
            M[i].copy_(enc["mask"])
            iters[i] = info["iters"]
        return run
    _, wall = run_batch([job(i) for i in range(N)], concurrency=conc, device=dev)
    return F, T, M, iters, wall


def train(net, F, T, M, tr, te, epochs, batch, lr0, warmup=100):
    """Adam over the training rows (shuffled per epoch); returns the per-epoch history."""
    import torch
    V = int(np.prod(DIMS))
    dev = F.device
    steps_per_epoch = len(tr) // batch
    total = epochs * steps_per_epoch
    rng = np.random.default_rng(0)
    hist = []
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        order = rng.permutation(tr)
        acc = torch.zeros((), dtype=torch.float64, device=dev)
        cnt = 0.0
        for s_ in range(steps_per_epoch):
            idx = torch.as_tensor(order[s_ * batch:(s_ + 1) * batch], device=dev)
            Fb = F.index_select(0, idx).view(-1)
            Tb = T.index_select(0, idx).view(-1)
            Mb = M.index_select(0, idx).view(-1)
            out = net.forward(Fb, DIMS, batch, keep=True)
            part, dout, c = net.loss(out, Tb, Mb, batch * V)
            net.zero_grad()
            net.backward(dout)
            lr = lr0 * min(1.0, (step + 1) / warmup) * 0.5 * (1 + math.cos(math.pi * step / max(1, total)))
            net.adam(max(lr, 1e-6))
            acc += part.sum()
            cnt += 3 * c
            step += 1
        train_mse = float(acc.item()) / cnt
        test_rel = evaluate_loss(net, F, T, M, te)
        hist.append(dict(epoch=ep + 1, train_mse=train_mse, test_rel_mse=test_rel, seconds=round(time.time() - t0, 1)))
        log(f"epoch {ep + 1}: train mse {train_mse:.5f}, test relative mse {test_rel:.4f} ({time.time() - t0:.0f} s)")
    return hist


def evaluate_loss(net, F, T, M, rows):
    """Relative masked MSE sum |pred - t|^2 / sum |t|^2 over the given rows."""
    import torch
    V = int(np.prod(DIMS))
    num = den = 0.0
    for i in rows:
        out = net.forward(F[i], DIMS, 1)
        m = M[i].bool()
        p = out.view(V, 16)[:, :3][m]
        t = T[i].view(V, 3)[m]
        num += float(((p - t) ** 2).sum().item())
        den += float((t ** 2).sum().item())
    return num / max(den, 1e-300)


def evaluate_solves(net, descs, dev, tol=1e-8, project=True):
    """Each held-out system solved plainly and from the network's prediction (one at a time,
    timed with the solver's own clock): iterations, solve ms, fallbacks, and the prediction's
    relative error against the solution."""
    import torch
    from comech_amd.batch import build_system, rhs
    from comech_amd.learned.data import encode, targets
    from comech_amd.learned.warmstart import learned_solve, predict
    rows = []
    for d in descs:
        s = build_system(d, device=dev)
        b = rhs(d, dev)
        x_p, ip = s.solve(b, tol=tol, maxiter=3000, mode="ref")
        enc = encode(d, dev)
        e0 = cell_rel = float("nan")
        if net is not None:
            x0 = predict(net, enc)
            e0 = float(torch.linalg.vector_norm(x0 - x_p) / torch.linalg.vector_norm(x_p))
            ts = float(net.meta["tscale"])
            V = int(np.prod(enc["dims"]))
            tc = targets(enc, x_p, ts).view(V, 3)
            pc = net.forward(enc["F"], enc["dims"], 1).view(V, 16)[:, :3]
            m = enc["mask"].bool()
            cell_rel = float(((pc - tc)[m] ** 2).sum() / (tc[m] ** 2).sum())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x_l, il = learned_solve(s, d, b, net, tol=tol, mode="ref", project=project, enc=enc)
        torch.cuda.synchronize()
        wall_l = (time.perf_counter() - t0) * 1e3
        # energy-norm error of the initial guess against the solution, relative to that of zero
        from comech_amd.learned.warmstart import CoarseSpace
        g0 = predict(net, enc) if net is not None else torch.zeros_like(x_p)
        if project:
            g0 = CoarseSpace(enc, g0).galerkin(s, b)
        eg = x_p - g0
        a_err = float(torch.sqrt(torch.dot(eg, s.apply(eg)) / torch.dot(x_p, s.apply(x_p))))
        rel = float(torch.linalg.vector_norm(x_l - x_p) / torch.linalg.vector_norm(x_p))
        true = float(torch.linalg.vector_norm(b - s.apply(x_l)) / torch.linalg.vector_norm(b))
        s.close()
        rows.append(dict(action=d["action"], plain_iters=ip["iters"], plain_ms=round(ip["t_solve_ms"], 2),
                         learned_iters=il["iters"], learned_solve_ms=round(il["t_solve_ms"], 2),
                         learned_wall_ms=round(wall_l, 2), fallback=il["fallback"], pred_rel_err=e0,
                         cell_rel_mse=cell_rel, guess_a_err=a_err,
                         r0_rel=il["r0_rel"], u_rel_vs_plain=rel, true_relres=true))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--feet", type=int, default=500, help="seeded feet (x 8 actions = solves)")
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--eval-feet", type=int, default=6, help="held-out feet solved with and without the warm start")
    ap.add_argument("--out", default="weights.npz")
    ap.add_argument("--summary", default=None)
    args = ap.parse_args()
    train_seeds, test_seeds = split_feet(args.feet)
    t0 = time.time()
    descs_tr = gather_inputs(train_seeds, args.workers)
    descs_te = gather_inputs(test_seeds, args.workers)
    log(f"inputs: {len(descs_tr)} training and {len(descs_te)} test systems in {time.time() - t0:.1f} s")
    import torch
    from comech_amd.learned.net import Net
    dev = torch.device("cuda:0")
    descs = descs_tr + descs_te
    F, T, M, iters, wall = solve_set(descs, dev)
    log(f"solved {len(descs)} systems in {wall:.1f} s ({len(descs) / wall:.1f} solves/s), "
        f"mean {np.mean(iters):.1f} iterations")
    ntr = len(descs_tr)
    tr, te = np.arange(ntr), np.arange(ntr, len(descs))
    V = int(np.prod(DIMS))
    ss = torch.zeros((), dtype=torch.float64, device=dev)
    na = 0.0
    for a in range(0, ntr, 64):                     # RMS of the active cells' targets, in chunks
        Tc = T[a:a + 64].view(-1, V, 3).double()
        Mc = M[a:a + 64].view(-1, V, 1).double()
        ss += (Tc * Tc * Mc).sum()
        na += 3.0 * float(Mc.sum().item())
    tscale = 1.0 / float(torch.sqrt(ss / na).item())
    T.mul_(tscale)
    log(f"target scale {tscale:.4g}")
    net = Net(dev, seed=1)
    hist = train(net, F, T, M, tr, te, args.epochs, args.batch, args.lr)
    net.save(args.out, tscale=tscale, dims=np.array(DIMS), h=H, feet=args.feet, epochs=args.epochs)
    del F, T, M
    torch.cuda.empty_cache()
    net.meta = dict(tscale=np.asarray(tscale))
    ev_descs = [d for d in descs_te[:8 * args.eval_feet]]
    rows = evaluate_solves(net, ev_descs, dev)
    rows_c = evaluate_solves(net, ev_descs, dev, project=False)
    rows_r = evaluate_solves(None, ev_descs, dev)
    summ = lambda rs, k: float(np.mean([r[k] for r in rs]))
    out = dict(dims=list(DIMS), h=H, feet=args.feet, train_systems=ntr, test_systems=len(descs_te),
               solve_wall_s=round(wall, 1), solve_iters_mean=float(np.mean(iters)), tscale=tscale, epochs=args.epochs,
               batch=args.batch, lr=args.lr, history=hist,
               eval=dict(systems=len(rows), plain_iters=summ(rows, "plain_iters"), learned_iters=summ(rows, "learned_iters"),
                         plain_ms=summ(rows, "plain_ms"), learned_wall_ms=summ(rows, "learned_wall_ms"),
                         learned_solve_ms=summ(rows, "learned_solve_ms"), fallbacks=int(sum(r["fallback"] for r in rows)),
                         pred_rel_err=summ(rows, "pred_rel_err"), cell_rel_mse=summ(rows, "cell_rel_mse"),
                         guess_a_err=summ(rows, "guess_a_err"), raw_guess_a_err=summ(rows_c, "guess_a_err"),
                         rigid_only_a_err=summ(rows_r, "guess_a_err"),
                         r0_rel=summ(rows, "r0_rel"),
                         raw_guess_iters=summ(rows_c, "learned_iters"), raw_guess_r0_rel=summ(rows_c, "r0_rel"),
                         raw_guess_wall_ms=summ(rows_c, "learned_wall_ms"),
                         rigid_only_iters=summ(rows_r, "learned_iters"), rigid_only_wall_ms=summ(rows_r, "learned_wall_ms"),
                         rows=rows, rows_raw_guess=rows_c, rows_rigid_only=rows_r))
    log(json.dumps({k: v for k, v in out["eval"].items() if not k.startswith("rows")}))
    if args.summary:
        with open(args.summary, "w") as fh:
            json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
