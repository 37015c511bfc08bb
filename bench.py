#!/usr/bin/env python3
"""Benchmark of the MI355X voxel FE solve path (one GPU).

Headline workload: the batched uneven-grid set -- 64 seeded hindfoot-like feet with grids drawn
uniformly between 96x64x64 and 192x112x112 voxels at 0.5 mm, each under 8 surgical actions
(baseline, resection, osteotomy, interposition, fusion at k = 1e6, loads x0.5/x1.5/x2), 512 solves
to relative residual 1e-8, every answer checked by its FP64 true residual.  One step is one foot
of the set under its 8 actions; ``--steps K`` feet are spread evenly over the 64 (K = 64 is the
whole set).  A step includes everything a solve needs after its inputs are on the device: the
multigrid hierarchy built on the GPU from the level-0 description (labels, split lists, spring
lists; the load actions reuse the baseline's hierarchy, the same operator), PCG and the check.
``--concurrency`` systems are in flight at once.

Prints ONE JSON line.  ``value`` = solves/s in mode ``ref`` (FP64 Krylov vectors, alpha/beta and
outer product, FP32 V-cycle: operator, smoothers and transfers, FP64 coarsest solve); also
``mixed`` (FP32 PCG inside FP64 iterative refinement) and ``plain_fp64`` (the reference's plain
method, all FP64: its level rule and per-node block Jacobi) on the same systems, batched, and
``plain_fp64_sequential`` (the plain method one solve at a time, as the reference runs it, on a
subset of the feet, with the batched method on the same subset), the error of every result
against the plain FP64 answer, ``roofline`` (the
FP32 fine-level K*u on the 384x256x256 grid, whose 635 MB working set streams from HBM) and
``cpu_baseline`` (the FP64 CPU oracle -- assembled CSR, Chebyshev GMG-PCG, SuperLU coarse solve,
the reference algorithm -- on one foot of the set, one core and a pool of host processes).
"""
import argparse
import glob
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_FEET = 64
# solves that share one system: the load actions have the baseline's operator
GROUPS = (("baseline", "load0.5", "load1.5", "load2"), ("resection",), ("osteotomy",), ("interposition",),
          ("fusion",))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    ram = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemTotal"):
                ram = round(int(line.split()[1]) / 1024 ** 2, 1)
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(nproc=os.cpu_count(), cores_available=avail, cpu_model=model, ram_gib=ram)


def pool(n):
    """Host worker processes (spawned; they import numpy/scipy only and never touch the GPU),
    single-threaded BLAS as the reference's production setting (exp/bank.py:10)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = "1"
    return ProcessPoolExecutor(max_workers=n, mp_context=mp.get_context("spawn"))


def timed_feet(k):
    """K feet spread evenly over the 64-foot set (all 64 for K = 64)."""
    return sorted({int(i * N_FEET // k) % N_FEET for i in range(k)}) if k <= N_FEET else \
        [i % N_FEET for i in range(k)]


# Multigrid of the method as built here: the semi-coarsened coarsest level and the split-copy
# smoother blocks (library defaults), coarsening down to 3,000 DOFs (7 levels at most).
TUNED = dict(min_dofs=3000, max_levels=7)
# The reference's plain method: its level rule (6,000 DOFs, 6 levels, every axis coarsened,
# fe/gmg.py:80-96) and per-node 3x3 block Jacobi (fe/gmg.py:29-44).
PLAIN = dict(last_axes=7, copy_blocks=False)


def foot_job(descs, bs, fi, grp, mode, tol, dev, ref_x=None, keep=False, mg=None):
    """One system (the first action of ``grp``) solved for every action of ``grp`` in turn."""
    import torch
    from comech_amd.batch import build_system
    from comech_amd.synth import ACTIONS
    idx = [ACTIONS.index(a) for a in grp]

    def run():
        s = build_system(descs[idx[0]], device=dev, **(mg or {}))
        ndof = s.levels()[0]["ndof"]
        out = []
        for k in idx:
            b = bs[k]
            x, info = s.solve(b, tol=tol, maxiter=3000, mode=mode)
            info["true_relres"] = float(torch.linalg.vector_norm(b - s.apply(x)) / torch.linalg.vector_norm(b))
            info.update(ndof=ndof, foot=fi, action=ACTIONS[k])
            if ref_x is not None:
                xr = ref_x[(fi, k)]
                info["u_rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(x - xr) / torch.linalg.vector_norm(xr))
            out.append(((fi, k), x if keep else None, info))
        s.close()
        return out
    return run


def batch_leg(feet, descs, bs, mode, tol, conc, dev, ref_x=None, keep=False, mg=None):
    """All solves of ``feet`` (8 each) batched; returns (infos, xs, wall seconds)."""
    from comech_amd.batch import run_batch
    jobs = []
    for fi in feet:
        for grp in GROUPS:
            w = len(grp) * descs[fi][0]["ncell"]
            jobs.append((w, foot_job(descs[fi], bs[fi], fi, grp, mode, tol, dev, ref_x, keep, mg)))
    jobs.sort(key=lambda j: -j[0])                # longest first: the tail of the batch stays short
    res, wall = run_batch([j[1] for j in jobs], concurrency=conc, device=dev)
    infos, xs = [], {}
    for out in res:
        for key, x, info in out:
            infos.append(info)
            if keep:
                xs[key] = x
    return infos, xs, wall


def lockstep_leg(feet, descs, bs, tol, width, dev, ref_x=None, mg=None):
    """All solves of ``feet`` in lock-step (comech_amd.batch.solve_lockstep): every launch of a PCG
    iteration covers all systems in the batch, systems join and leave as they finish; the load
    actions reuse their baseline's system.  Returns (infos, wall seconds)."""
    import torch
    from comech_amd.batch import build_system, solve_lockstep
    from comech_amd.synth import ACTIONS
    items = []
    for fi in feet:
        for grp in GROUPS:
            idx = [ACTIONS.index(a) for a in grp]

            def mk(fi=fi, idx=idx):
                s = build_system(descs[fi][idx[0]], device=dev, **(mg or {}))
                return s, [((fi, k), bs[fi][k]) for k in idx]
            items.append((len(idx) * descs[fi][0]["ncell"], mk))
    items.sort(key=lambda j: -j[0])
    got, wall = solve_lockstep([j[1] for j in items], width=width, builders=4, tol=tol, device=dev)
    infos = []
    for (fi, k), (x, info) in got.items():
        b = bs[fi][k]
        info["ndof"] = int(b.numel())
        info.update(foot=fi, action=ACTIONS[k])
        if ref_x is not None:
            xr = ref_x[(fi, k)]
            info["u_rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(x - xr) / torch.linalg.vector_norm(xr))
        infos.append(info)
    return infos, wall


def summarize(infos, wall, extra=None):
    import numpy as np
    iters = [i["iters"] for i in infos]
    d = {"value": round(len(infos) / wall, 3), "unit": "solves/s", "solves": len(infos),
         "wall_s": round(wall, 3), "iters_mean": round(float(np.mean(iters)), 1), "iters_max": int(max(iters)),
         "true_relres_max": max(i["true_relres"] for i in infos),
         "dof_iterations_per_s": round(sum(i["ndof"] * i["iters"] for i in infos) / wall, 1)}
    if extra:
        d.update(extra)
    return d


SAMPLE_SCALE = 2     # the CPU sample: one foot of the set at twice the voxel size (1/8 of the cells)


def oracle_sample(foot, tol):
    """CPU worker: the FP64 oracle on one foot of the set (baseline action) drawn at twice the voxel
    size.  Returns the solve's own wall interval (generation excluded), iterations, displacement
    and von Mises."""
    import numpy as np
    from comech_amd import synth
    from oracle import fe_oracle as O
    c = synth.batch_case(foot, "baseline", scale=SAMPLE_SCALE)
    g = c["grid"]
    og = O.Grid(g.xe, g.ye, g.ze)
    t0 = time.time()
    U, info, R = O.solve_cn(og, c["lam"], c["mu"], c["active"], c["cn"], c["orig"], c["fixed"], c["loads"],
                            extra=c["extra_cb"], tol=tol, maxiter=3000)
    t1 = time.time()
    act = np.flatnonzero(c["active"])
    _, vm, _ = O.cell_stress(og, c["lam"], c["mu"], U, act, cn=c["cn"])
    return dict(t0=t0, t1=t1, iters=int(info["iters"]), ndof=int(info["ndof"]), U=U, vm=vm, ncell=g.ncell)


def cpu_baseline(ex, nworkers, sample_foot, tol, cells_per_solve, gpu_answer):
    """The oracle (reference algorithm, FP64) on one foot of the set: alone on one core (latency),
    then ``nworkers`` copies at once on the host cores (throughput, the reference's process pool,
    exp/rb_run_engine_v4.py:123); solves/s extrapolated to the timed set by cell count."""
    import numpy as np
    one = ex.submit(oracle_sample, sample_foot, tol).result()
    dt1 = one["t1"] - one["t0"]
    many = list(ex.map(oracle_sample, [sample_foot] * nworkers, [tol] * nworkers))
    t0, t1 = min(m["t0"] for m in many), max(m["t1"] for m in many)
    scale = cells_per_solve / one["ncell"]
    pool_rate = nworkers / (t1 - t0) / scale
    info = host_info()
    out = {"value": round(pool_rate, 5), "unit": "solves/s", "cores": nworkers, "kind": "port",
           "sample": (f"foot {sample_foot} of the set at 1 mm voxels ({one['ncell']} cells, {one['ndof']} DOFs, "
                      f"baseline action; the set is at 0.5 mm): the FP64 oracle (assembled CSR, Galerkin GMG-PCG with Chebyshev/block-Jacobi, SuperLU "
                      f"coarse solve) to rtol {tol:g}, {one['iters']} iterations; {nworkers} single-threaded processes "
                      f"solving it at once took {t1 - t0:.1f} s, extrapolated to the timed set's mean of "
                      f"{cells_per_solve:.0f} cells per solve (x{scale:.2f} in cell count)"),
           "single_core": {"value": round(1.0 / (dt1 * scale), 5), "unit": "solves/s", "solve_s": round(dt1, 2)}}
    out.update(info)
    if gpu_answer is not None:
        U, vm = gpu_answer
        Uo = np.nan_to_num(one["U"])
        out["gpu_vs_oracle"] = {"u_rel_l2": float(np.linalg.norm(np.nan_to_num(U) - Uo) / np.linalg.norm(Uo)),
                                "vm_rel_l2": float(np.linalg.norm(vm - one["vm"]) / np.linalg.norm(one["vm"])),
                                "what": "mode ref on the GPU vs the oracle, both to the same rtol, same system"}
    return out


def roofline_case():
    """CPU worker: level-0 description of the 384x256x256 grid at 0.25 mm (4-voxel shells)."""
    from comech_amd import synth
    return synth.level0_description(synth.hindfoot_large())


def roofline(desc, dev, nrep=20):
    """Fine-level kernels on the HBM-resident grid, timed with HIP events on the system's stream:
    K*u (FP32 and FP64) and one fused Chebyshev smoothing step (FP32), at the z extents the
    project's autotuner picks for this grid on the idle GPU (products and smoothing steps tuned
    separately)."""
    import torch
    from comech_amd.batch import build_system
    s = build_system(desc, device=dev)
    table = tuning_table(s)
    if table is None:
        s.autotune(("f32", "f64"), nrep=5)
    else:
        s.set_tuning(table)
    st = s.level_stats(0)
    kz = {k: st[k] for k in ("kz_f32", "kz_f32_smooth", "kz_f64")}
    nodes = s.grid.nnode
    ncell = desc["ncell"]
    u32 = torch.randn(3 * s.nn, dtype=torch.float32, device=dev)
    y32 = torch.empty_like(u32)
    # the smoothing step first: its set-up (block-Jacobi inverses, power iteration) launches K*u, and
    # the timed launches of each kernel are then its last ones in a kernel trace
    mss = s.time_apply(u32, y32, nrep=nrep, smooth_step=True)
    ms32 = s.time_apply(u32, y32, nrep=nrep, structured_only=True)
    ms64 = s.time_apply(u32.double(), y32.double(), nrep=nrep, structured_only=True)
    s.close()
    alg = 2 * 3 * 4 * nodes + ncell                      # read u, write y (FP32) + one label byte per cell
    alg64 = 2 * 3 * 8 * nodes + ncell
    algs = 24 * 4 * nodes + ncell                        # s, rin, Dinv (6), x in; r, d, x out (FP32)
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
    return dict(nodes=nodes, cells=ncell, kz=kz, ms32=ms32, ms64=ms64, mss=mss, alg=alg, alg64=alg64, algs=algs,
                gbs32=gbs(alg, ms32), gbs64=gbs(alg64, ms64), gbss=gbs(algs, mss))


TUNING = os.path.join(ROOT, "comech_amd", "tuning.json")


def grid_key(s):
    lv = s.levels()[0]
    return f"{lv['nx']}x{lv['ny']}x{lv['nz']}"


def tuning_table(s):
    """The stored autotuner table for this system's grid (comech_amd/tuning.json), or None."""
    try:
        return json.load(open(TUNING))["grids"].get(grid_key(s))
    except (OSError, ValueError, KeyError):
        return None


def pmc_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of ``kernel``."""
    for p in sorted(glob.glob(os.path.join(ROOT, "profiles", "pmc_*.json"))):
        try:
            d = json.load(open(p))
        except Exception:
            continue
        if d.get("kernel") == kernel:
            return d.get("traffic_bytes")
    return None


def batch_workload(args):
    import numpy as np
    from comech_amd import synth
    feet = timed_feet(args.steps)
    warm = [f for f in reversed(range(N_FEET)) if f not in feet][:args.warmup] or feet[:args.warmup]
    nw = max(1, min(16, host_info()["cores_available"] or 1))
    ex = pool(nw)
    t0 = time.time()
    fut_large = ex.submit(roofline_case)              # the first submission starts the host workers
    from comech_amd.learned import train as learned_train
    learned_seeds = learned_train.split_feet(500)[1][:LEARNED_FEET]
    fut_learned = [ex.submit(learned_train.foot_inputs, sd) for sd in learned_seeds]
    futs = {f: ex.submit(synth.batch_foot, f) for f in sorted(set(feet) | set(warm))}
    import torch
    from comech_amd.batch import build_system, rhs
    torch.cuda.init()
    dev = torch.device("cuda:0")
    descs = {f: fu.result() for f, fu in futs.items()}
    bs = {f: [rhs(d, dev) for d in descs[f]] for f in descs}
    torch.cuda.synchronize()
    sizes = {f: tuple(len(e) - 1 for e in descs[f][0]["edges"]) for f in descs}
    log(f"inputs of {len(descs)} feet ready in {time.time() - t0:.1f} s on {nw} host workers")
    conc = args.concurrency
    if warm:
        batch_leg(warm, descs, bs, "ref", args.tol, conc, dev, mg=TUNED)
        batch_leg(warm, descs, bs, "fp64", args.tol, conc, dev, mg=PLAIN)
        log("warmup done")
    plain, plain_x, w_plain = batch_leg(feet, descs, bs, "fp64", args.tol, conc, dev, keep=True, mg=PLAIN)
    log(f"plain fp64 (reference method): {len(plain) / w_plain:.2f} solves/s batched")
    ref, _, w_ref = batch_leg(feet, descs, bs, "ref", args.tol, conc, dev, ref_x=plain_x, mg=TUNED)
    log(f"ref: {len(ref) / w_ref:.2f} solves/s")
    mixed, _, w_mixed = batch_leg(feet, descs, bs, "mixed", args.tol, conc, dev, ref_x=plain_x, mg=TUNED)
    log(f"mixed: {len(mixed) / w_mixed:.2f} solves/s")
    lock, w_lock = lockstep_leg(feet, descs, bs, args.tol, LOCK_WIDTH, dev, ref_x=plain_x, mg=TUNED)
    for i in lock:
        i["true_relres"] = i["relres"]
    log(f"lock-step: {len(lock) / w_lock:.2f} solves/s")
    del plain_x
    seq, _, w_seq = batch_leg(feet[:1], descs, bs, "ref", args.tol, 1, dev, mg=TUNED)
    # the plain method as the reference runs it, one solve at a time, on a subset of the timed feet,
    # and the batched method on the same subset
    sub = feet[::max(1, len(feet) // SEQ_FEET)][:SEQ_FEET]
    pseq, _, w_pseq = batch_leg(sub, descs, bs, "fp64", args.tol, 1, dev, mg=PLAIN)
    rsub, _, w_rsub = batch_leg(sub, descs, bs, "ref", args.tol, conc, dev, mg=TUNED)
    log(f"plain fp64 one at a time: {len(pseq) / w_pseq:.2f} solves/s; ref batched on the same feet: "
        f"{len(rsub) / w_rsub:.2f}")
    bad = [dict(foot=i["foot"], action=i["action"], mode=i["mode"], flag=i["flag"], true=i["true_relres"])
           for i in plain + ref + mixed + pseq + rsub + lock if i["flag"] or i["true_relres"] > args.tol * 1.0001]
    for b in bad:
        log("missed", b)
    if any(b["flag"] == 0 for b in bad) or any(i["flag"] for i in ref):
        log(f"bench.py: a solve missed rtol {args.tol}")
        sys.exit(1)
    ll = learned_leg([d for fu in fut_learned for d in fu.result()], dev, args.tol)
    log(f"learned: {ll['learned_iters']} vs {ll['plain_iters']} iterations, {ll['learned_ms']} vs {ll['plain_ms']} ms")
    rl = roofline(fut_large.result(), dev)
    log(f"roofline: K*u FP32 {rl['ms32'] * 1e3:.1f} us, smoothing step {rl['mss'] * 1e3:.1f} us")
    cells = float(np.mean([descs[f][0]["ncell"] for f in feet]))
    # the oracle's sample: the smallest foot of the timed set, baseline action; the GPU's answer to
    # the same system in mode ref for the accuracy against the oracle
    sample = min(feet, key=lambda f: descs[f][0]["ncell"])
    cpu = None
    if not args.no_cpu_baseline:
        from comech_amd.grid import Grid, corners_of
        from comech_amd.system import cell_stress_gpu
        d0 = ex.submit(synth.batch_foot, sample, scale=SAMPLE_SCALE, actions=["baseline"]).result()[0]
        s = build_system(d0, device=dev, **TUNED)
        x, _ = s.solve(rhs(d0, dev), tol=args.tol, maxiter=3000, mode="ref")
        U = x.view(-1, 3).cpu().numpy().copy()
        U[~s.used_mask(0)] = np.nan
        s.close()
        g = Grid(*d0["edges"])
        lab = np.asarray(d0["labels"])
        act = np.flatnonzero(lab != 255)
        nodes8 = corners_of(g, act)
        pos = np.searchsorted(act, d0["irr_cells"])
        nodes8[pos] = d0["irr_nodes"]
        lam = np.where(lab != 255, np.asarray(d0["lam_tab"])[np.minimum(lab, len(d0["lam_tab"]) - 1)], 0.0)
        mu = np.where(lab != 255, np.asarray(d0["mu_tab"])[np.minimum(lab, len(d0["mu_tab"]) - 1)], 0.0)
        vm = cell_stress_gpu(g, lam, mu, x, act, nodes8=nodes8, device=dev)[1]
        cpu = cpu_baseline(ex, nw, sample, args.tol, cells, (U, vm))
        log(f"cpu baseline: {cpu['value']} solves/s on {nw} cores, one core {cpu['single_core']['value']}")
    ex.shutdown()
    nsolve = len(ref)
    err = lambda infos: [round(i["u_rel_l2_vs_fp64"], 15) for i in infos]
    kname = "k_march<float, LabelMat<float,SRI,ISO>, EPI_APPLY> (fine-level K*u)"
    line = {
        "metric": "solves/s to rtol 1e-8 (batched) + matvec HBM GB/s; learned/mixed vs plain",
        "value": round(nsolve / w_ref, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * w_ref / len(feet), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"batched uneven grids: {len(feet)} of the 64 seeded feet (grids 96x64x64 to 192x112x112 "
                                f"voxels at 0.5 mm, each its own seed) x 8 surgical actions = {nsolve} solves to rtol "
                                f"{args.tol:g}, each checked by its FP64 true residual; one step = one foot's 8 "
                                f"actions (hierarchy built on the GPU per system, the load actions reuse the "
                                f"baseline's); mode ref: FP64 PCG with an FP32 V-cycle"),
                   "feet": feet, "grid_min": [int(min(sizes[f][a] for f in feet)) for a in range(3)],
                   "grid_max": [int(max(sizes[f][a] for f in feet)) for a in range(3)], "solves": nsolve,
                   "concurrency": conc,
                   "mode": "ref", **{k: v for k, v in summarize(ref, w_ref).items() if k not in ("value", "unit")}},
        "mixed": summarize(mixed, w_mixed, {"ratio_over_plain": round(w_plain / w_mixed, 3)}),
        "lockstep": summarize(lock, w_lock, {
            "what": f"mode ref in lock-step: {LOCK_WIDTH} systems at a time, every launch of a PCG iteration covering "
                    f"all of them (cm_batch_*), systems joining and leaving as they finish",
            "lockstep_solves": int(sum(bool(i.get("lockstep")) for i in lock)),
            "ref_max_vs_fp64": max(i["u_rel_l2_vs_fp64"] for i in lock)}),
        "plain_fp64": summarize(plain, w_plain, {
            "what": "the reference's plain method in FP64 on the GPU (its level rule: 6,000 DOFs or 6 levels, every "
                    "axis coarsened; per-node 3x3 block Jacobi; Chebyshev degree 3), batched like the headline"}),
        "plain_fp64_sequential": summarize(pseq, w_pseq, {
            "what": f"the same plain method one solve at a time (as the reference runs it, one process per GPU; the "
                    f"residual checked every iteration, as in every leg), feet {sub} x 8 actions",
            "ref_batched_same_feet": round(len(rsub) / w_rsub, 3),
            "ratio_ref_batched_over_plain_sequential": round((len(rsub) / w_rsub) / (len(pseq) / w_pseq), 3)}),
        "ratio_ref_over_plain": round(w_plain / w_ref, 3),
        "multigrid": {"ref_and_mixed": dict(TUNED, last_axes=3, copy_blocks=True),
                      "plain_fp64": dict(PLAIN, min_dofs=6000, max_levels=6)},
        "learned": ll,
        "sequential": {"value": round(len(seq) / w_seq, 3), "unit": "solves/s", "concurrency": 1,
                       "solves": len(seq), "what": f"foot {feet[0]}'s 8 actions one at a time, mode ref"},
        "accuracy": {"what": "displacement relative L2 of every result against the plain FP64 GMG-PCG answer of the "
                             "same system (itself pinned to the FP64 CPU oracle by the parity tests)",
                     "ref_max": max(err(ref)), "mixed_max": max(err(mixed)), "ref": err(ref), "mixed": err(mixed),
                     "order": [f"{i['foot']}:{i['action']}" for i in ref]},
        "roofline": {"bound": "hbm", "achieved": round(rl["gbs32"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(rl["gbs32"] / HBM_PEAK_GBS, 4), "traffic": pmc_traffic("fine_apply_f32_384x256x256"),
                     "kernel": kname, "workload": "384x256x256 voxels at 0.25 mm (25.2 M cells, 25.4 M nodes): "
                                                  "working set 635 MB, above the 256 MB Infinity Cache",
                     "avg_launch_us": round(rl["ms32"] * 1e3, 2), "alg_bytes_per_launch": rl["alg"],
                     "alg_bytes": "24 B per node (read u, write y, FP32) + 1 B label per cell",
                     "z_extents": rl["kz"]},
        "roofline_smoothing_step": {"kernel": "k_march<float, LabelMat<float,SRI,ISO>, EPI_CHEB> (fused Chebyshev "
                                              "step, the solve's dominant kernel)",
                                    "achieved": round(rl["gbss"], 1), "frac": round(rl["gbss"] / HBM_PEAK_GBS, 4),
                                    "avg_launch_us": round(rl["mss"] * 1e3, 2), "alg_bytes_per_launch": rl["algs"],
                                    "traffic": pmc_traffic("fine_cheb_f32_384x256x256"),
                                    "alg_bytes": "96 B per node (u, rin, Dinv 6, x in; r, d, x out; FP32) + 1 B/cell"},
        "roofline_fp64": {"achieved": round(rl["gbs64"], 1), "frac": round(rl["gbs64"] / HBM_PEAK_GBS, 4),
                          "avg_launch_us": round(rl["ms64"] * 1e3, 2), "alg_bytes_per_launch": rl["alg64"]},
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


LEARNED_FEET = 2      # held-out feet of the learned warm start's training split, 8 actions each
SEQ_FEET = 8          # feet of the one-at-a-time plain leg
LOCK_WIDTH = 16       # systems in one lock-step batch


def learned_leg(descs, dev, tol):
    """The learned warm start on feet the network never saw (96x64x64, the training family): every
    system solved from zero and from the learned guess (prediction projected onto its coarse
    space), one at a time, wall-clocked around the whole call (prediction, projection, solve)."""
    import numpy as np
    import torch
    from comech_amd.batch import build_system, rhs
    from comech_amd.learned import load_default
    from comech_amd.learned.data import encode
    from comech_amd.learned.warmstart import learned_solve
    net = load_default(dev)
    rows = []
    for i, d in enumerate([descs[0]] + descs):       # the first system once more, untimed (warm-up)
        s = build_system(d, device=dev)
        b = rhs(d, dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xp, ip = s.solve(b, tol=tol, maxiter=3000, mode="ref")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        enc = encode(d, dev)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        xl, il = learned_solve(s, d, b, net, tol=tol, mode="ref", enc=enc)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        true = float(torch.linalg.vector_norm(b - s.apply(xl)) / torch.linalg.vector_norm(b))
        s.close()
        if i:
            rows.append(dict(action=d["action"], plain_iters=ip["iters"], learned_iters=il["iters"],
                             plain_ms=(t1 - t0) * 1e3, learned_ms=(t3 - t2) * 1e3, guess_ms=il["t_guess_ms"],
                             fallback=il["fallback"], true_relres=true))
    m = lambda k: round(float(np.mean([r[k] for r in rows])), 2)
    return {"what": f"{len(rows)} held-out systems (96x64x64 voxels, 8 actions each), solved one at a time in mode "
                    f"ref from zero and from the learned initial guess; wall ms per solve (learned: input encoding "
                    f"excluded, prediction and projection included)",
            "plain_iters": m("plain_iters"), "learned_iters": m("learned_iters"), "plain_ms": m("plain_ms"),
            "learned_ms": m("learned_ms"), "guess_ms": m("guess_ms"),
            "ratio_learned_over_plain": round(m("plain_ms") / max(m("learned_ms"), 1e-9), 3),
            "fallbacks": int(sum(r["fallback"] for r in rows)),
            "true_relres_max": max(r["true_relres"] for r in rows)}


def retune(desc, dev, runs=3):
    """Autotuner table of the grid of ``desc`` into comech_amd/tuning.json: ``runs`` autotuner runs
    on fresh systems, each entry the extent most runs picked (ties: the shorter chain)."""
    from collections import Counter
    from comech_amd.batch import build_system
    tabs = []
    for r in range(runs):
        s = build_system(desc, device=dev)
        s.autotune(("f32", "f64"), nrep=5, retime=True)
        tabs.append(s.tuning())
        key = grid_key(s)
        s.close()
        log(f"retune run {r}: {tabs[-1]}")
    table = [row[:3] + [sorted(Counter(t[l][q] for t in tabs).items(), key=lambda kv: (-kv[1], kv[0]))[0][0]
                        for q in range(3, 7)] for l, row in enumerate(tabs[0])]
    d = json.load(open(TUNING))
    d["grids"][key] = table
    write_tuning(d)
    log(f"tuning table for {key}: {table}")


def write_tuning(d):
    """comech_amd/tuning.json with one level per line."""
    grids = ",\n".join(f'  "{k}": [\n' + ",\n".join("   " + json.dumps(r) for r in v) + "\n  ]"
                        for k, v in d["grids"].items())
    with open(TUNING, "w") as f:
        f.write("{\n" + f' "what": {json.dumps(d["what"])},\n "how": {json.dumps(d["how"])},\n'
                + ' "grids": {\n' + grids + "\n }\n}\n")


def kernel_workload(args):
    """For rocprofv3 runs (kernel trace, or one --pmc counter set per run): the 384x256x256 system
    built in this process (no host workers), the fine-level kernels timed as in the headline run,
    then 5 plain device copies of 256 MiB (counter calibration).  Prints the roofline figures."""
    import torch
    from comech_amd import synth
    torch.cuda.init()
    dev = torch.device("cuda:0")
    t0 = time.time()
    desc = synth.level0_description(synth.hindfoot_large())
    log(f"384x256x256 inputs ready in {time.time() - t0:.1f} s")
    if args.retune:
        retune(desc, dev)
    rl = roofline(desc, dev, nrep=args.nrep)
    a = torch.randn(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "fine-level kernels, 384x256x256", **rl}), flush=True)


def solves_workload(args):
    """For rocprofv3 kernel traces of the solve mix: the batched ref leg alone on ``--steps`` feet of
    the set with ``--concurrency`` systems in flight, after one warm-up foot; ``--eager`` launches
    the V-cycles without HIP graphs (every launch visible to the trace).  Inputs are built in this
    process (no host workers).  Prints one JSON line (not the headline metric)."""
    import torch
    from comech_amd import synth
    import comech_amd.batch as CB
    feet = timed_feet(args.steps)
    warm = [f for f in reversed(range(N_FEET)) if f not in feet][:1]
    descs = {f: synth.batch_foot(f) for f in feet + warm}
    if args.eager:
        build = CB.build_system

        def build_eager(*a, **k):
            s = build(*a, **k)
            s.set_graphs(False)
            return s
        CB.build_system = build_eager
    torch.cuda.init()
    dev = torch.device("cuda:0")
    bs = {f: [CB.rhs(d, dev) for d in descs[f]] for f in descs}
    log(f"inputs of {len(descs)} feet ready")
    batch_leg(warm, descs, bs, "ref", args.tol, args.concurrency, dev, mg=TUNED)
    infos, _, wall = batch_leg(feet, descs, bs, "ref", args.tol, args.concurrency, dev, mg=TUNED)
    print(json.dumps({"workload": f"batched set, feet {feet}, mode ref, {args.concurrency} in flight"
                                  + (", eager V-cycles" if args.eager else ""), **summarize(infos, wall)}), flush=True)


def large_workload(args):
    """Large single solve (the fifth configuration): 384x256x256 voxels at 0.25 mm, 4-voxel cortical
    shells, one solve per precision form.  Prints one JSON line (not the headline metric): fine
    matvec time and algorithmic GB/s per precision, V-cycle time, solve time and iterations to rtol
    1e-8 (FP64 true residual) for fp64 / ref / mixed, and the von Mises relative L2 of the mixed
    forms against the FP64 answer (ml/evaluate.py:93,96 metric)."""
    import numpy as np
    import torch
    from comech_amd import synth
    from comech_amd.engine import split_cells
    from comech_amd.system import VoxelSystem, cell_stress_gpu
    dev = torch.device("cuda:0")
    t0 = time.time()
    c = synth.hindfoot_large()
    extra, terms, f = synth.solve_inputs_hindfoot(c)
    irr_c, irr_n = split_cells(c["grid"], c["active"], c["cn"])
    g = c["grid"]
    log(f"large inputs ready in {time.time() - t0:.1f} s: ncell {g.ncell}, active {int(c['active'].sum())}")
    f_dev = torch.from_numpy(f).to(dev)
    s = VoxelSystem(g, sri=True, active=c["active"], labels=c["labels"], lam_tab=c["lam_tab"], mu_tab=c["mu_tab"],
                    copy_orig=c["orig"][g.nnode:], irr_cells=irr_c, irr_nodes=irr_n, extra=extra, rank1=terms,
                    device=dev, levels=args.levels, last_axes=args.last_axes)
    table = tuning_table(s)                 # the autotuner's z extents for this grid (or tune now:
    if table is None:                        # one large solve has the GPU to itself)
        s.autotune(("f32", "f64"), nrep=5)
    else:
        s.set_tuning(table)
    lv = s.levels()
    out = {"workload": "hindfoot raster 384x256x256 @0.25mm (4-voxel cortical shells), rtol 1e-8, FP64 true residual",
           "ncell": g.ncell, "dofs": lv[0]["ndof"], "levels": [l["ndof"] for l in lv], "last_axes": args.last_axes,
           "t_setup_ms": round(s.timings()["total"], 1)}
    for dt, name, sz in ((torch.float32, "f32", 4), (torch.float64, "f64", 8)):
        u = torch.randn(3 * s.nn, dtype=dt, device=dev)
        y = torch.empty_like(u)
        ms = s.time_apply(u, y, nrep=20, structured_only=True)
        alg = 2 * 3 * sz * g.nnode + g.ncell
        out[f"matvec_{name}"] = {"us": round(ms * 1e3, 1), "alg_bytes": alg, "GBps": round(alg / (ms * 1e-3) / 1e9, 1),
                                 "frac_of_8TBs": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        r = torch.randn(3 * s.nn, dtype=dt, device=dev)
        z = torch.empty_like(r)
        s.vcycle(r, z)
        torch.cuda.synchronize()
        t1 = time.time()
        for _ in range(5):
            s.vcycle(r, z)
        torch.cuda.synchronize()
        out[f"vcycle_{name}_ms"] = round((time.time() - t1) / 5 * 1e3, 2)
        del u, y, r, z
    sols = {}
    for mode in ("fp64", "ref", "mixed"):
        x, info = s.solve(f_dev, tol=args.tol, maxiter=3000, mode=mode)
        rr = float(torch.linalg.vector_norm(f_dev - s.apply(x)) / torch.linalg.vector_norm(f_dev))
        out[mode] = {"iters": info["iters"], "flag": info["flag"], "t_solve_ms": round(info["t_solve_ms"], 1),
                     "true_relres": rr}
        sols[mode] = x
        log(mode, out[mode])
    cells = np.flatnonzero(c["active"])
    nodes8 = c["cn"][cells]
    vm = {m: cell_stress_gpu(g, c["lam"], c["mu"], sols[m], cells, nodes8=nodes8)[1] for m in sols}
    for m in ("ref", "mixed"):
        out[m]["vm_rel_l2_vs_fp64"] = float(np.linalg.norm(vm[m] - vm["fp64"]) / np.linalg.norm(vm["fp64"]))
        out[m]["u_rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(sols[m] - sols["fp64"]) /
                                           torch.linalg.vector_norm(sols["fp64"]))
    s.close()
    print(json.dumps(out), flush=True)



def prep_workload(args):
    """Engine preparation at the hindfoot size: a 27-sample lattice built from the raster (bone
    bodies of the engine's label table, coalition bar, soft tissue; 30 % of cells mix 9 samples
    with their +x neighbour), then lattice materials + bone shares + parts and split connectivity.
    GPU (cm_lattice_materials, cm_split_connectivity, results copied back) against the oracle's
    restatement of the same reference algorithm on the host (one thread).  One JSON line."""
    import numpy as np
    import torch
    from comech_amd import synth
    from comech_amd.rb import rbfe as RF
    from comech_amd.rb import splitfe as SF
    os.environ["OMP_NUM_THREADS"] = "1"
    dev = torch.device("cuda:0")
    c = synth.hindfoot_raster(dims=(160, 96, 96), h=0.5)
    g = c["grid"]
    lab = c["labels"]
    bodies = [("calcaneus", 0), ("talus", 0), ("navicular", 0)]
    body_of = np.zeros(256, np.int64)
    body_of[0], body_of[1], body_of[255] = 61, RF.COAL, 0
    for b, (name, _) in enumerate(bodies):
        body_of[2 + 2 * b] = RF.BONES[name][0]
        body_of[3 + 2 * b] = RF.BONES[name][1]
    cls = {61: 5, RF.COAL: 1}
    for name, _ in bodies:
        cls[RF.BONES[name][0]], cls[RF.BONES[name][1]] = 1, 2
    base = body_of[lab]
    rng = np.random.default_rng(7)
    L = np.repeat(base[:, None], 27, 1).astype(np.int16)
    nbr = np.roll(base.reshape(g.nx, g.ny, g.nz), -1, 0).ravel()
    mix = rng.uniform(size=g.ncell) < 0.3
    L[mix, :9] = nbr[mix, None]
    V = dict(L=L, cls=cls, xe=g.xe, ye=g.ye, ze=g.ze)
    for _ in range(2):                                     # warm up (library load, caches)
        o = RF.lattice_prep(V, device=dev, host=False)
        SF.split_connectivity_gpu(g, o["active"], o["part"], o["S"], device=dev)
    torch.cuda.synchronize()
    t0 = time.time()
    nrep = 5
    for _ in range(nrep):
        o = RF.lattice_prep(V, device=dev, host=False)
        cn, orig, ns = SF.split_connectivity_gpu(g, o["active"], o["part"], o["S"], device=dev)
        E = o["E"].cpu().numpy()
    torch.cuda.synchronize()
    t_gpu = (time.time() - t0) / nrep
    log(f"prep: lattice + split on the GPU {t_gpu * 1e3:.1f} ms")
    line = {"workload": "engine preparation, hindfoot raster 160x96x96 as a 27-sample lattice",
            "ncell": g.ncell, "samples": int(L.size), "split_nodes": ns, "copies": int(len(orig) - g.nnode),
            "gpu_ms": round(t_gpu * 1e3, 2)}
    if not args.no_cpu_baseline:
        from oracle import fe_oracle as O
        t1 = time.time()
        Eo, nuo, acto, Fo, sero = O.lattice_materials(L, cls, RF.ROLES, (g.nx, g.ny, g.nz))
        So, _ = O.bone_share(L, cls, RF.ROLES)
        po = O.lattice_part(So, Fo)
        og = O.Grid(g.xe, g.ye, g.ze)
        cno, origo, nso = O.split_connectivity_fast(og, acto, po, So)
        t_cpu = time.time() - t1
        log(f"prep: lattice + split on the host {t_cpu:.2f} s")
        line["cpu_baseline"] = {"seconds": round(t_cpu, 2), "cores": 1, "kind": "port",
                                "sample": "the same lattice, oracle restatement of rb/rbfe.py:75-141 and "
                                          "rb/splitfe.py:261-305 on the host"}
        line["speedup"] = round(t_cpu / t_gpu, 1)
        line["parity"] = bool(np.array_equal(cn, cno) and np.array_equal(orig, origo) and ns == nso
                              and np.allclose(E, Eo, rtol=1e-12, atol=0))
    # preparation around the solve (rb/rbfe.py:166-178): node sets of every part, the bone of each
    # end of 40 ligaments (seeded points in the bones' boxes), the ligament patches and matrix
    cn_h, orig_h = cn, orig
    act, part, names = o["active"].cpu().numpy(), o["part"].cpu().numpy(), o["names"]
    keys = sorted(names)
    rng = np.random.default_rng(11)
    sets, _ = RF.part_node_sets(cn_h, part, act, len(orig_h), keys, device=dev)
    bone_nodes = {names[k]: sets[k] for k in keys if len(sets[k])}
    xyz_of = lambda nodes: np.stack([g.xe[orig_h[nodes] % (g.nx + 1)], g.ye[(orig_h[nodes] // (g.nx + 1)) % (g.ny + 1)],
                                     g.ze[orig_h[nodes] // ((g.nx + 1) * (g.ny + 1))]], 1)
    springs = []
    for _ in range(40):
        ends = []
        for nm in rng.choice(list(bone_nodes), 2):
            P = xyz_of(bone_nodes[nm])
            ends.append((P.min(0) + rng.uniform(size=3) * (P.max(0) - P.min(0))).tolist())
        springs.append({"name": "", "k": 100.0, "a": ends[0], "b": ends[1]})
    used_all = np.ones(len(orig_h), bool)

    def gpu_solve_prep():
        st, _ = RF.part_node_sets(cn_h, part, act, len(orig_h), keys, device=dev)
        bn = {names[k]: st[k] for k in keys if len(st[k])}
        sp_ = [dict(x) for x in springs]
        for x, (ea, eb) in zip(sp_, RF.spring_end_bones(sp_, bn, g, orig_h, device=dev)):
            x["A"], x["B"] = ea, eb
        return bn, sp_, RF.spring_matrix_cn(g, orig_h, sp_, bn, device=dev).rank1(used_all)
    gpu_solve_prep()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(nrep):
        bn, sp_, terms = gpu_solve_prep()
    torch.cuda.synchronize()
    t_sp = (time.time() - t0) / nrep
    log(f"prep: solve preparation on the GPU {t_sp * 1e3:.1f} ms")
    line["solve_prep"] = {"what": "node sets of every bone part, nearest-bone ends of 40 ligaments, ligament patches "
                                  "and their rank-one weights (rb/rbfe.py:166-178, rb/replica3.py:68-86)",
                          "gpu_ms": round(t_sp * 1e3, 2), "ligament_weights": int(len(terms[1])),
                          "patch_nodes_mean": round(len(terms[1]) / 6.0 / len(springs), 1)}
    if not args.no_cpu_baseline:
        from scipy.spatial import cKDTree
        t1 = time.time()
        cells = np.arange(g.ncell)
        so = {k: np.unique(cn_h[cells[(part == k) & act]].ravel()) for k in keys}
        bno = {names[k]: so[k] for k in keys if len(so[k])}
        xyz = xyz_of(np.arange(len(orig_h)))
        trees = {nm: cKDTree(xyz[nd]) for nm, nd in bno.items()}
        spo = [dict(x) for x in springs]
        for x in spo:
            for key, p in (("A", x["a"]), ("B", x["b"])):
                x[key] = min(trees, key=lambda nm: trees[nm].query(p)[0])
        pat = O.spring_patches(xyz, spo, bno, used_all)
        t_cpu_sp = time.time() - t1
        log(f"prep: solve preparation on the host {t_cpu_sp:.2f} s")
        line["solve_prep"]["cpu_baseline"] = {"seconds": round(t_cpu_sp, 3), "cores": 1, "kind": "port",
                                              "sample": "the same parts and springs: np.unique per part, one cKDTree "
                                                        "per bone (the reference's method), the oracle's patch search "
                                                        "(distances and a sort per end)"}
        line["solve_prep"]["speedup"] = round(t_cpu_sp / t_sp, 1)
        pat_g = RF.spring_patches(sp_, bn, g, orig_h, used=used_all, device=dev)
        line["solve_prep"]["parity"] = bool(all(np.array_equal(bn[nm], bno[nm]) for nm in bno)
                                            and [(x["A"], x["B"]) for x in sp_] == [(x["A"], x["B"]) for x in spo]
                                            and all(np.array_equal(a, c) and np.array_equal(b_, d)
                                                    for (a, b_), (c, d) in zip(pat_g, pat)))
    # analysis-grid resampling (rb/rbsim.py:110-161) onto the engine's 60x40x44 grid of 2 mm cells
    # (rb/rbsim.py:15-17), centred on the raster: a smooth displacement field on the FE nodes (NaN on
    # unused nodes), random cell stresses; connectivity and displacement resident on the GPU
    from comech_amd.grid import Grid
    from comech_amd.rb import rbsim as RS
    mid = [(e[0] + e[-1]) / 2 for e in (g.xe, g.ye, g.ze)]
    gm = Grid(*[m - 1.0 * n + 2.0 * np.arange(n + 1) for m, n in zip(mid, (60, 40, 44))])
    nxyz = xyz_of(np.arange(len(orig_h)))
    Urel = np.stack([np.sin(nxyz[:, 0] / 9.0), np.cos(nxyz[:, 1] / 7.0), nxyz[:, 2] / 50.0], 1) * 0.1
    Urel[np.setdiff1d(np.arange(len(orig_h)), np.unique(cn_h[act]))] = np.nan
    VM = rng.uniform(0, 30, g.ncell).astype(np.float32)
    SIG = rng.standard_normal((g.ncell, 6)).astype(np.float32)
    o_dev = dict(g=g, cn=torch.from_numpy(cn_h).to(dev), active=torch.from_numpy(act).to(dev),
                 VM=torch.from_numpy(VM).to(dev), SIG=torch.from_numpy(SIG).to(dev))
    U_dev = torch.from_numpy(Urel).to(dev)

    def gpu_resample():
        vm_c, u_c, cov = RS.sample_fe(o_dev, gm, U_dev, sub=2, device=dev)
        return vm_c, u_c, cov, RS.sample_nodes(o_dev, gm, U_dev, device=dev)
    gpu_resample()
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(nrep):
        res_g = gpu_resample()
    torch.cuda.synchronize()
    t_rs = (time.time() - t0) / nrep
    log(f"prep: resampling on the GPU {t_rs * 1e3:.1f} ms")
    line["resample"] = {"what": "sample_fe (sub 2) and sample_nodes onto the 60x40x44 analysis grid at 2 mm "
                                "(rb/rbsim.py:110-161); results copied to the host",
                        "gpu_ms": round(t_rs * 1e3, 2), "analysis_cells": int(gm.ncell),
                        "analysis_nodes": int(gm.nnode)}
    if not args.no_cpu_baseline:
        t1 = time.time()
        og = O.Grid(g.xe, g.ye, g.ze)
        vmo, uo, covo, _ = O.sample_fe(og, cn_h, act, VM, SIG, Urel, gm, sub=2)
        uno = O.sample_nodes(og, cn_h, act, Urel, gm)
        t_cpu_rs = time.time() - t1
        log(f"prep: resampling on the host {t_cpu_rs:.2f} s")
        line["resample"]["cpu_baseline"] = {"seconds": round(t_cpu_rs, 3), "cores": 1, "kind": "port",
                                            "sample": "the same fields: the oracle's restatement of rb/rbsim.py:110-161 "
                                                      "(numpy, scipy distance_transform_edt)"}
        line["resample"]["speedup"] = round(t_cpu_rs / t_rs, 1)
        vm_g, u_g, cov_g, un_g = res_g
        line["resample"]["parity"] = bool(np.array_equal(un_g, uno) and np.array_equal(cov_g, covo)
                                          and np.abs(vm_g - vmo).max() <= 1e-5 * np.abs(vmo).max()
                                          and np.abs(u_g - uo).max() <= 1e-5 * np.abs(uo).max())
    print(json.dumps(line), flush=True)




def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=N_FEET,
                    help="feet of the 64-foot set in the timed region (8 solves each), spread evenly; 64 = all")
    ap.add_argument("--warmup", type=int, default=2, help="untimed feet before the timed region")
    ap.add_argument("--concurrency", type=int, default=16, help="systems in flight on the GPU")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="batch", choices=["batch", "kernels", "solves", "large", "prep"],
                    help="batch: the headline (uneven-grid set); kernels: the fine-level kernels of the "
                         "384x256x256 grid alone, for rocprofv3; solves: the batched ref leg alone, for "
                         "rocprofv3; large: the 384x256x256 solve in every mode; prep: engine preparation "
                         "and resampling (separate JSON lines, not the headline)")
    ap.add_argument("--eager", action="store_true", help="solves: V-cycles launched without HIP graphs")
    ap.add_argument("--nrep", type=int, default=20, help="kernels: launches timed per kernel")
    ap.add_argument("--retune", action="store_true",
                    help="kernels: rerun the autotuner on the 384x256x256 grid and rewrite its table in "
                         "comech_amd/tuning.json before timing")
    ap.add_argument("--levels", type=int, default=None, help="large: force the number of multigrid levels")
    ap.add_argument("--last-axes", type=int, default=3,
                    help="large: axes the step to the coarsest level coarsens (x 1, y 2, z 4; 7 = fe/gmg.py's rule)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; every solve is independent (batched solves "
              "shard by solve across GPUs), so run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    {"batch": batch_workload, "kernels": kernel_workload, "solves": solves_workload, "large": large_workload,
     "prep": prep_workload}[
        args.workload](args)


if __name__ == "__main__":
    main()
