"""Batched solves on one GPU: many independent systems (feet x surgical actions, uneven grids) in
flight at once.

The reference runs one solve per process and round-robins processes over GPUs
(rb/gpusolve.py:174-180, exp/rb_run_engine_v4.py:19-23).  Here one process keeps several solves in
flight on one GPU: each worker thread owns a HIP stream (default priority), builds its system on
that stream and solves it on the system's own stream with the captured V-cycle graph.  A single
solve cannot fill 256 CUs on its small coarse levels (a few thousand cells, latency-bound
launches); independent solves overlap those phases with each other's fine-level work.  The C ABI
releases the GIL in every call, so the threads run their host work concurrently too.
"""
import queue
import threading
import time

import torch

from .system import require_gpu


def run_batch(jobs, concurrency=4, device=None):
    """Run ``jobs`` (callables without arguments, each building and solving one system and
    returning its result) on ``concurrency`` worker threads.  Returns (results in job order,
    wall seconds).  The first exception raised by a job is re-raised after all workers stop."""
    dev = require_gpu(device)
    jobs = list(jobs)
    results = [None] * len(jobs)
    errors = []
    q = queue.Queue()
    for i, fn in enumerate(jobs):
        q.put((i, fn))
    nw = max(1, min(int(concurrency), len(jobs)))
    for _ in range(nw):
        q.put(None)

    def worker():
        torch.cuda.set_device(dev)
        st = torch.cuda.Stream(device=dev)
        with torch.cuda.stream(st):
            while True:
                item = q.get()
                if item is None:
                    break
                i, fn = item
                try:
                    results[i] = fn()
                except BaseException as e:      # noqa: BLE001 -- reported to the caller below
                    errors.append((i, e))
        st.synchronize()

    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker, daemon=True) for _ in range(nw)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0][1]
    return results, wall


def build_system(desc, device=None, **mg):
    """VoxelSystem of a level-0 description (``synth.level0_description``): labels + tables, split
    copies and irregular cells, ground springs (COO) and spring terms (rank one)."""
    import numpy as np
    from .grid import Grid
    from .system import VoxelSystem
    g = Grid(*desc["edges"])
    return VoxelSystem(g, sri=True, active=np.asarray(desc["labels"]) != 255, labels=desc["labels"],
                       lam_tab=desc["lam_tab"], mu_tab=desc["mu_tab"], copy_orig=desc["copy_orig"],
                       irr_cells=desc["irr_cells"], irr_nodes=desc["irr_nodes"], extra=desc["extra"],
                       rank1=desc["terms"], device=device, **mg)


def rhs(desc, device=None):
    """Nodal force vector of a level-0 description on the device (FP64, 3 per node)."""
    dev = require_gpu(device)
    b = torch.zeros(3 * desc["nn"], dtype=torch.float64, device=dev)
    b[torch.as_tensor(desc["f_idx"], device=dev)] = torch.as_tensor(desc["f_val"], dtype=torch.float64, device=dev)
    return b


__all__ = ["run_batch", "build_system", "rhs"]
