#!/usr/bin/env python3
"""Benchmark of the MI355X voxel FE solve path (one GPU).

Workload: the synthetic hindfoot raster (160x96x96 voxels at 0.5 mm, ~2.38M active DOFs, split
nodes with interface springs, 2000 ground springs, 200 ligament springs, 700 N load) under the 8
surgical actions (baseline, resection, osteotomy, interposition, fusion, load x0.5/x1.5/x2), each
solved to relative residual 1e-8 with mixed-precision GMG-PCG: FP32 operator, smoothers, transfers
and V-cycle; FP64 Krylov vectors, alpha/beta reductions and dense coarsest solve; every answer
checked by an FP64 true residual.  One step = one complete solve (GPU multigrid setup from the
level-0 description, PCG, check); ``--concurrency`` solves are in flight at once (batched).

Prints ONE JSON line.  ``value`` = solves/s (batched); ``sequential`` = the same solves one at a
time; ``plain_fp64`` = all-FP64 GMG-PCG batched the same way; ``roofline`` = the FP32 fine-level
stencil kernel (the matvec) against the 8 TB/s HBM peak; ``cpu_baseline`` = the FP64 CPU oracle
(assembled CSR, Chebyshev GMG-PCG, SuperLU coarse solve -- the reference algorithm) on a bounded
sample.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_inputs(dims=(160, 96, 96), h=0.5, shell=2):
    from comech_amd import synth
    from comech_amd.engine import split_cells
    c = synth.hindfoot_raster(dims=dims, h=h, shell=shell)
    extra, terms, f = synth.solve_inputs_hindfoot(c)
    irr_c, irr_n = split_cells(c["grid"], c["active"], c["cn"])
    return c, (extra, terms), f, irr_c, irr_n


def make_system(c, extra, irr_c, irr_n, kz_cache=None):
    from comech_amd.system import VoxelSystem
    s = VoxelSystem(c["grid"], sri=True, active=c["active"], labels=c["labels"], lam_tab=c["lam_tab"],
                    mu_tab=c["mu_tab"], copy_orig=c["orig"][c["grid"].nnode:], irr_cells=irr_c, irr_nodes=irr_n,
                    extra=extra[0], rank1=extra[1], device="cuda:0", smoother="chebyshev")
    s.autotune()          # structured-block z extents: timed once per level shape, then cached
    return s


def one_solve(c, extra, f_dev, irr_c, irr_n, mode, tol):
    import torch
    s = make_system(c, extra, irr_c, irr_n)
    x, info = s.solve(f_dev, tol=tol, maxiter=2000, mode=mode)
    info["ndof"] = s.levels()[0]["ndof"]
    r = f_dev - s.apply(x)                               # FP64 true residual
    relres = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(f_dev))
    s.close()
    info["true_relres"] = relres
    return info


def cpu_baseline(sample_dims=(80, 48, 48)):
    """Oracle (reference algorithm, FP64, one thread) on a bounded sample of the same anatomy;
    solves/s extrapolated linearly in the number of cells to the full raster."""
    os.environ["OMP_NUM_THREADS"] = "1"
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    from comech_amd import synth
    from oracle import fe_oracle as O
    c = synth.hindfoot_raster(dims=sample_dims, h=80.0 / sample_dims[0])
    g = O.Grid(c["grid"].xe, c["grid"].ye, c["grid"].ze)
    t0 = time.time()
    U, info, R = O.solve_cn(g, c["lam"], c["mu"], c["active"], c["cn"], c["orig"], c["fixed"], c["loads"],
                            extra=c["extra_cb"], tol=1e-8, maxiter=2000)
    dt = time.time() - t0
    scale = (160 * 96 * 96) / float(g.ncell)
    return dict(value=1.0 / (dt * scale), unit="solves/s", cores=1, kind="port",
                sample=(f"hindfoot anatomy at {sample_dims[0]}x{sample_dims[1]}x{sample_dims[2]} cells "
                        f"({g.ncell} cells, {info['ndof']} DOFs): one FP64 oracle solve (assembly, Galerkin "
                        f"hierarchy, Chebyshev GMG-PCG, SuperLU coarse) to rtol 1e-8 in {dt:.2f} s, "
                        f"{info['iters']} iterations, extrapolated x{scale:.0f} in cell count"))


def pmc_traffic(kernel="k_march"):
    """HBM bytes per launch of the fine stencil from the committed rocprofv3 --pmc summary."""
    for p in sorted(glob.glob(os.path.join(ROOT, "profiles", "pmc_*.json"))):
        try:
            d = json.load(open(p))
        except Exception:
            continue
        if d.get("kernel", "").startswith(kernel):
            return d.get("traffic_bytes")
    return None


def action_inputs(actions):
    """Inputs of the hindfoot raster under each surgical action (same anatomy seed)."""
    from comech_amd import synth
    from comech_amd.engine import split_cells
    out = []
    for a in actions:
        c = synth.hindfoot_raster(dims=(160, 96, 96), h=0.5, action=a)
        extra, terms, f = synth.solve_inputs_hindfoot(c)
        irr_c, irr_n = split_cells(c["grid"], c["active"], c["cn"])
        out.append((c, (extra, terms), f, irr_c, irr_n))
    return out


def large_workload(args):
    """Large single solve: 384x256x256 voxels at 0.25 mm (the hindfoot anatomy upscaled, 4-voxel cortical
    shells), one solve per precision form.  Prints one JSON line (not the headline metric): fine
    matvec time and algorithmic GB/s per precision, V-cycle time, solve time and iterations to rtol
    1e-8 (FP64 true residual) for fp64 / ref / mixed, and the von Mises relative L2 of the mixed
    forms against the FP64 answer (ml/evaluate.py:93,96 metric)."""
    import numpy as np
    import torch
    from comech_amd.system import cell_stress_gpu
    dev = torch.device("cuda:0")
    t0 = time.time()
    c, extra, f, irr_c, irr_n = build_inputs(dims=(384, 256, 256), h=0.25, shell=4)
    g = c["grid"]
    log(f"large inputs ready in {time.time() - t0:.1f} s: ncell {g.ncell}, active {int(c['active'].sum())}")
    f_dev = torch.from_numpy(f).to(dev)
    s = make_system(c, extra, irr_c, irr_n)
    lv = s.levels()
    out = {"workload": "hindfoot raster 384x256x256 @0.25mm (the hindfoot anatomy upscaled, 4-voxel cortical shells), rtol 1e-8, FP64 true residual",
           "ncell": g.ncell, "dofs": lv[0]["ndof"], "levels": [l["ndof"] for l in lv],
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


def batch3_workload(args):
    """Uneven-grid batch: the first ``--feet`` feet of the 64-foot set (grids drawn uniformly
    between 96x64x64 and 192x112x112 at 0.5 mm, each with its own seed) under the 8 surgical
    actions, solved batched (``--concurrency`` in flight) to rtol 1e-8 in mode ``--mode``, every
    answer checked by the FP64 true residual.  One solve per grid shape runs alone first (the
    z-extent autotuner times its kernels on an idle GPU).  Prints one JSON line (not the headline
    metric): solves/s, DOF-iterations/s, iterations, and the fine K*u FP32 rate of the smallest and
    the largest grid."""
    import numpy as np
    import torch
    from comech_amd import synth
    from comech_amd.batch import run_batch
    from comech_amd.engine import split_cells
    from comech_amd.synth import ACTIONS
    dev = torch.device("cuda:0")
    t0 = time.time()
    sizes = synth.batch_sizes(64)[:args.feet]
    cases = []
    for foot in range(args.feet):
        for a in ACTIONS:
            c = synth.batch_case(foot, a)
            extra, terms, f = synth.solve_inputs_hindfoot(c)
            irr_c, irr_n = split_cells(c["grid"], c["active"], c["cn"])
            cases.append((c, (extra, terms), torch.from_numpy(f).to(dev), irr_c, irr_n))
        log(f"batch3: foot {foot} {sizes[foot]} inputs ready ({time.time() - t0:.1f} s)")

    def job(i):
        c, extra, f_dev, irr_c, irr_n = cases[i]
        return lambda: one_solve(c, extra, f_dev, irr_c, irr_n, args.mode, args.tol)
    for foot in range(args.feet):                      # tune every shape alone
        run_batch([job(foot * len(ACTIONS))], concurrency=1, device=dev)
    log("batch3: shapes tuned")
    infos, wall = run_batch([job(i) for i in range(len(cases))], concurrency=args.concurrency, device=dev)
    worst = max(i["true_relres"] for i in infos)
    missed = [dict(foot=i // len(ACTIONS), action=ACTIONS[i % len(ACTIONS)], iters=r["iters"], flag=r["flag"],
                   outer=r["outer"], relres=r["relres"], true_relres=r["true_relres"])
              for i, r in enumerate(infos) if r["flag"] or r["true_relres"] > args.tol * 1.0001]
    for m in missed:
        log("batch3: missed", m)
    rate = {}
    for name, foot in (("smallest", int(np.argmin([np.prod(z) for z in sizes]))),
                       ("largest", int(np.argmax([np.prod(z) for z in sizes])))):
        c, extra, f_dev, irr_c, irr_n = cases[foot * len(ACTIONS)]
        sy = make_system(c, extra, irr_c, irr_n)
        u = torch.randn(3 * sy.nn, dtype=torch.float32, device=dev)
        y = torch.empty_like(u)
        ms = sy.time_apply(u, y, nrep=20, structured_only=True)
        g = c["grid"]
        alg = 2 * 3 * 4 * g.nnode + g.ncell
        rate[name] = {"grid": list(sizes[foot]), "us": round(ms * 1e3, 2), "alg_bytes": alg,
                      "GBps": round(alg / (ms * 1e-3) / 1e9, 1)}
        sy.close()
    iters = [i["iters"] for i in infos]
    line = {"workload": f"uneven-grid batch: {args.feet} feet (grids 96x64x64 .. 192x112x112 at 0.5 mm) x "
                        f"{len(ACTIONS)} surgical actions, rtol {args.tol}, FP64 true residual",
            "mode": args.mode, "concurrency": args.concurrency, "solves": len(infos),
            "grids": [list(z) for z in sizes], "solves_per_s": round(len(infos) / wall, 2),
            "dof_iterations_per_s": round(sum(i["ndof"] * i["iters"] for i in infos) / wall, 1),
            "iters_mean": round(float(np.mean(iters)), 1), "iters_max": int(max(iters)),
            "true_relres_max": worst, "missed": missed, "matvec_f32": rate}
    print(json.dumps(line), flush=True)
    # a solve the solver reports as converged must meet the tolerance; a solve it reports as not
    # converged (flag != 0: restarts from the true residual stopped improving, the attainable FP64
    # accuracy of that system lies above the tolerance) is listed above, not hidden
    if any(m["flag"] == 0 for m in missed):
        log(f"bench.py: a solve reported converged but missed rtol {args.tol}")
        sys.exit(1)


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
    ap.add_argument("--steps", type=int, default=24, help="solves in the timed region (the surgical actions, cycled)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes over the action set")
    ap.add_argument("--concurrency", type=int, default=16, help="solves in flight on the GPU")
    ap.add_argument("--mode", default="ref", choices=["mixed", "ref", "fp64"])
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="hindfoot", choices=["hindfoot", "large", "prep", "batch3"],
                    help="large: the 384x256x256 single solve; prep: engine preparation and resampling; "
                         "batch3: uneven grids of the 64-foot batch (separate JSON lines, not the headline metric)")
    ap.add_argument("--feet", type=int, default=8, help="batch3: feet of the 64-foot set (x 8 actions)")
    ap.add_argument("--kernel-only", action="store_true",
                    help="for rocprofv3 --pmc runs: 10 launches of the FP32 fine stencil on the workload, then "
                         "5 plain device copies of 256 MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; every solve is independent (batched solves "
              "shard by solve across GPUs), so run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from comech_amd.batch import run_batch
    from comech_amd.synth import ACTIONS
    torch.cuda.init()
    dev = torch.device("cuda:0")
    t0 = time.time()
    if args.workload == "large":
        large_workload(args)
        return
    if args.workload == "prep":
        prep_workload(args)
        return
    if args.workload == "batch3":
        batch3_workload(args)
        return
    if args.kernel_only:
        c, extra, f, irr_c, irr_n = build_inputs()
        s = make_system(c, extra, irr_c, irr_n)
        u = torch.randn(3 * s.nn, dtype=torch.float32, device=dev)
        y = torch.empty_like(u)
        for _ in range(5):
            s.time_apply(u, y, nrep=1, structured_only=True)          # 2 launches each
        a = torch.randn(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        s.close()
        log("kernel-only probe done")
        return

    inputs = action_inputs(ACTIONS)
    f_devs = [torch.from_numpy(i[2]).to(dev) for i in inputs]
    c0 = inputs[0][0]
    log(f"inputs ready in {time.time() - t0:.1f} s: {len(inputs)} actions, ncell {c0['grid'].ncell}, "
        f"active {int(c0['active'].sum())}")

    def job(i, mode):
        c, extra, f, irr_c, irr_n = inputs[i % len(inputs)]
        return lambda: one_solve(c, extra, f_devs[i % len(inputs)], irr_c, irr_n, mode, args.tol)

    def timed(mode, n, conc):
        res, wall = run_batch([job(i, mode) for i in range(n)], concurrency=conc, device=dev)
        return res, wall

    # the first solve runs alone: the z-extent autotuner (cached per level shape) must time its
    # kernels on an otherwise idle GPU, not between the kernels of concurrent solves
    timed(args.mode, 1, 1)
    for _ in range(args.warmup):
        res, _ = timed(args.mode, len(inputs), args.concurrency)
        log("warmup", [round(r["t_solve_ms"], 1) for r in res])
    infos, dt = timed(args.mode, args.steps, args.concurrency)
    for i in infos:
        log("step", {k: i[k] for k in ("iters", "flag", "outer", "levels", "relres", "true_relres", "t_setup_ms",
                                       "t_solve_ms", "mode")})
    worst = max(i["true_relres"] for i in infos)
    if worst > args.tol * 1.0001 or any(i["flag"] for i in infos):
        log(f"bench.py: a solve missed rtol {args.tol}: {worst}")
        sys.exit(1)
    # the same jobs one at a time, and plain FP64 GMG-PCG (the reference method) batched the same way
    seq, dt_seq = timed(args.mode, len(inputs), 1)
    plain, dt_plain = timed("fp64", args.steps, args.concurrency)
    sps = args.steps / dt
    plain_sps = args.steps / dt_plain
    iters = [i["iters"] for i in infos]
    dof_it = sum(i["ndof"] * i["iters"] for i in infos) / dt

    # roofline of the dominant kernel: FP32 fine-level stencil (HIP events on the solver's stream)
    c, extra, f, irr_c, irr_n = inputs[0]
    s = make_system(c, extra, irr_c, irr_n)
    g = c["grid"]
    u = torch.randn(3 * s.nn, dtype=torch.float32, device=dev)
    y = torch.empty_like(u)
    nrep = 50
    ms = s.time_apply(u, y, nrep=nrep, structured_only=True)
    ms64 = s.time_apply(u.double(), y.double(), nrep=nrep, structured_only=True)
    s.close()
    alg_bytes = 2 * 3 * 4 * g.nnode + g.ncell             # read u, write y (FP32) + one label byte per cell
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    line = {
        "metric": "solves/s to rtol 1e-8 (batched) + matvec HBM GB/s; learned/mixed vs plain",
        "value": round(sps, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "hindfoot raster 160x96x96 @0.5mm under the 8 surgical actions, mixed-precision "
                               "GMG-PCG to rtol 1e-8 checked by an FP64 true residual; one step = one solve "
                               "(GPU multigrid setup + PCG + check)",
                   "dofs": infos[0]["ndof"], "mode": args.mode, "concurrency": args.concurrency,
                   "iters_mean": round(float(np.mean(iters)), 1), "levels": infos[-1]["levels"],
                   "t_setup_ms": round(float(np.mean([i["t_setup_ms"] for i in infos])), 2),
                   "t_solve_ms": round(float(np.mean([i["t_solve_ms"] for i in infos])), 2),
                   "true_relres_max": worst,
                   "dof_iterations_per_s": round(dof_it, 1)},
        "sequential": {"value": round(len(inputs) / dt_seq, 4), "unit": "solves/s", "concurrency": 1},
        "plain_fp64": {"value": round(plain_sps, 4), "unit": "solves/s", "concurrency": args.concurrency,
                       "iters_mean": round(float(np.mean([i["iters"] for i in plain])), 1),
                       "ratio_mixed_over_plain": round(sps / plain_sps, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(),
                     "kernel": "k_march<float, LabelMat<float,SRI,ISO>, APPLY> (fine-level K*u)",
                     "avg_launch_us": round(ms * 1e3, 2),
                     "alg_bytes_per_launch": alg_bytes, "fp64_avg_launch_us": round(ms64 * 1e3, 2)},
    }
    if not args.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline()
        except Exception as e:          # the baseline is reported, never required
            line["cpu_baseline"] = {"value": None, "error": repr(e)[:200]}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
