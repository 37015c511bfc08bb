"""Drop-in for the reference's ``rb/gpusolve.py`` GPU solver entry points (``pcg_gmg_gpu``,
``pick_device``, ``free_gpu``, ``GPUMG``; ``hierarchy`` and ``assemble_gpu`` raise: the GPU path forms
no CSR matrices).

``pcg_gmg_gpu(A, b, used_nodes, grid_edges, tol, maxiter, x0, Z, device, mg_dtype)`` keeps the
reference signature (rb/gpusolve.py:265-292): FP64 PCG with the V-cycle in ``mg_dtype`` (FP32 by
default) and the residual checked at iterations < 5 and every 5th (rb/gpusolve.py:244).  A failed
attempt is retried up to four times after releasing cached memory, as the reference does; unlike
the reference there is no CPU fallback afterwards -- the last error is raised.  The test hooks of
the reference are honoured before any device call: ``COMECH_TEST_GPU_FAIL`` (every attempt fails)
and ``COMECH_TEST_GPU_FAIL_ONCE=1`` (the first attempt fails).  ``COMECH_RETRY_SLEEP`` scales the
back-off sleeps (default 1; the reference sleeps 5 + 10 a + U(0,10) s).
"""
import gc
import os
import random
import time

import torch

from ..fe import gmg as _gmg

os.environ.setdefault("PYTORCH_NVML_BASED_CUDA_CHECK", "1")


def pick_device():
    """rb/gpusolve.py:174-180: COMECH_GPU or pid % device_count; None without a GPU."""
    if not torch.cuda.is_available():
        return None
    dev = os.environ.get("COMECH_GPU")
    if dev is None:
        dev = str(os.getpid() % torch.cuda.device_count())
    return f"cuda:{dev}"


def free_gpu(dev=None):
    """rb/gpusolve.py:192-203."""
    gc.collect()
    if torch.cuda.is_available():
        try:
            if dev is not None and str(dev).startswith("cuda"):
                with torch.cuda.device(dev):
                    torch.cuda.empty_cache()
            else:
                torch.cuda.empty_cache()
        except Exception:
            pass


def hierarchy(A, used_nodes, xe, ye, ze, min_dofs=6000, max_levels=6):
    """rb/gpusolve.py:39-65 returns the Galerkin levels as host scipy CSR matrices and prolongations.
    The GPU path builds its hierarchy on the device (matrix-free fine level, exact moment cells on
    coarse levels, a dense coarsest inverse) and never forms CSR levels, so this raises.  The level
    table of the same rule is ``GPUMG.from_csr(A, used_nodes, (xe, ye, ze)).levels()``."""
    raise NotImplementedError("hierarchy: the GPU solver forms no CSR levels; use GPUMG.from_csr(...).levels()")


def assemble_gpu(g, lam, mu, cells, cn, KLf, dev, chunk=60000):
    """rb/gpusolve.py:306-347 assembles the global CSR matrix on the GPU for the CSR solver.  The
    matrix-free GPU path applies the element operator directly (comech_amd.rb.splitfe.solve_cn)
    and has no CSR assembly, so this raises."""
    raise NotImplementedError("assemble_gpu: the GPU path is matrix-free; call comech_amd.rb.splitfe.solve_cn")


class GPUMG:
    """The multigrid preconditioner object of rb/gpusolve.py:132-171 on the GPU hierarchy.

    Build it with ``GPUMG.from_csr(A, used_nodes, grid_edges, device, dt)`` (the inputs of
    pcg_gmg); ``vcycle(b)`` applies one V-cycle to a device vector in the compact DOF order of A
    (dtype float32 or float64 as ``dt``); ``mode`` names the coarsest-level method.  The reference
    constructor ``GPUMG(As, Ps, dev, dt)`` takes host CSR levels, which this path does not form
    (see ``hierarchy``): called that way it raises."""
    mode = "dense-inverse"

    def __init__(self, As=None, Ps=None, dev=None, dt=None):
        raise NotImplementedError("GPUMG(As, Ps, ...) needs host CSR levels; build it with GPUMG.from_csr(A, used_nodes, "
                                  "grid_edges, device, dt)")

    @classmethod
    def from_csr(cls, A, used_nodes, grid_edges, device=None, dt=torch.float32, **mg):
        import numpy as np
        import scipy.sparse as sp
        self = cls.__new__(cls)
        self.dev = torch.device(device or pick_device())
        self.dt = dt
        self.sys, node_of = _gmg.csr_system(sp.csr_matrix(A), used_nodes, grid_edges, device=self.dev, **mg)
        self.dofs = torch.as_tensor((3 * node_of[:, None] + np.arange(3)).ravel(), device=self.dev)
        self.levels = self.sys.levels
        return self

    def vcycle(self, b):
        full = torch.zeros(3 * self.sys.nn, dtype=self.dt, device=self.dev)
        full[self.dofs] = b.to(self.dt)
        return self.sys.vcycle(full)[self.dofs]

    def close(self):
        self.sys.close()


def _attempt(A, b, used_nodes, grid_edges, tol, maxiter, x0, Z, dev, mode, **opts):
    if os.environ.get("COMECH_TEST_GPU_FAIL"):
        raise RuntimeError("test: simulated GPU failure")
    if os.environ.get("COMECH_TEST_GPU_FAIL_ONCE") == "1":
        os.environ["COMECH_TEST_GPU_FAIL_ONCE"] = "0"
        raise RuntimeError("test: simulated transient GPU failure")
    return _gmg.pcg_gmg(A, b, used_nodes, grid_edges, tol=tol, maxiter=maxiter, x0=x0, Z=Z, device=dev, mode=mode,
                        check_every=opts.pop("check_every", 5), **opts)


def pcg_gmg_gpu(A, b, used_nodes, grid_edges, tol=1e-7, maxiter=800, x0=None, Z=None, device=None,
                mg_dtype="float32", **opts):
    dev = device or pick_device()
    if dev is None:
        raise RuntimeError("pcg_gmg_gpu: no GPU visible (the MI355X solver has no CPU fallback)")
    mode = opts.pop("mode", "ref" if mg_dtype == "float32" else "fp64")
    scale = float(os.environ.get("COMECH_RETRY_SLEEP", "1"))
    err = None
    for attempt in range(4):
        try:
            x, info = _attempt(A, b, used_nodes, grid_edges, tol, maxiter, x0, Z, dev, mode, **dict(opts))
            info.update(gpu_attempt=attempt)
            return x, info
        except Exception as e:                       # noqa: BLE001 (retry any device failure)
            err = e
        free_gpu(dev)
        if os.environ.get("COMECH_TEST_GPU_FAIL"):
            break
        time.sleep(scale * (5 + 10 * attempt + 10 * random.random()))
    raise RuntimeError(f"pcg_gmg_gpu failed after retries: {err!r}") from err
