"""Host wrapper of one GPU voxel system (ctypes over libcomech_hip.so).

A ``VoxelSystem`` owns the device hierarchy of one problem: the fine level is matrix-free
(per-cell labels or materials, split copies, irregular cells, sparse extra stiffness), coarse levels
hold exact Galerkin moment cells, the coarsest level a dense FP64 inverse.  Vectors are torch CUDA
tensors in the level-0 node layout (3 values per node, grid nodes in the reference's numbering
followed by split copies).  PyTorch only provides device memory and the stream.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import P_double, P_ll, P_u8, check, ptr

MODES = {"fp64": 0, "ref": 1, "mixed": 2}
SMOOTHERS = {"chebyshev": 0, "jacobi": 1}
PREC = {"f64": 0, "fp64": 0, torch.float64: 0, "f32": 1, "fp32": 1, torch.float32: 1}


def require_gpu(device=None):
    if not torch.cuda.is_available():
        raise RuntimeError("comech_amd: no HIP device visible; the solver runs on the GPU only (no CPU fallback)")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError(f"comech_amd: device {dev} is not a GPU")
    return dev


def compress_materials(lam, mu, active, max_labels=255):
    """Per-cell (lam, mu) -> (table_lam, table_mu, labels) when the active cells use at most
    ``max_labels`` distinct pairs (exact: the table holds the cells' own values), else None."""
    act = np.asarray(active, bool)
    la, ma = np.asarray(lam, float)[act], np.asarray(mu, float)[act]
    ul, il = np.unique(la, return_inverse=True)
    if len(ul) > max_labels:
        return None
    um, im = np.unique(ma, return_inverse=True)
    if len(um) > max_labels:
        return None
    code = il.astype(np.int64) * len(um) + im
    uc, ic = np.unique(code, return_inverse=True)
    if len(uc) > max_labels:
        return None
    lab = np.full(len(act), 255, np.uint8)
    lab[act] = ic.astype(np.uint8)
    return ul[uc // len(um)], um[uc % len(um)], lab


class VoxelSystem:
    def __init__(self, grid, *, sri, active, lam=None, mu=None, labels=None, lam_tab=None, mu_tab=None,
                 copy_orig=None, irr_cells=None, irr_nodes=None, extra=None, rank1=None, used=None, min_dofs=6000,
                 max_levels=6,
                 levels=None, smoother="chebyshev", degree=3, omega=0.6, npre=2, npost=2, materials="auto",
                 last_axes=3, copy_blocks=True, march_blocks=600, device=None):
        """grid: Grid; active: (ncell,) bool in reference cell order; materials either per cell
        (lam, mu) or as labels + tables; split connectivity as copy originals and irregular cells
        with their 8 node ids; extra: (rows, cols, vals) COO over level-0 DOFs.  last_axes: axes
        (x 1, y 2, z 4) the step to the coarsest level coarsens; 7 is the reference rule of
        fe/gmg.py, the default 3 keeps the vertical resolution on the coarsest grid.  copy_blocks:
        level-0 smoother blocks over each grid node and its split copies when extra stiffness couples
        them (False: the per-node 3x3 block Jacobi of fe/gmg.py:29-44).  march_blocks: the z-extent
        heuristic takes the longest z chain that still gives this many march blocks per launch (fewer
        suit solves that share the GPU with others)."""
        self.device = require_gpu(device)
        self.lib = _lib.load()
        self.grid = grid
        self.sri = bool(sri)
        self.stream = torch.cuda.current_stream(self.device)
        self.nng = grid.nnode
        self.ncopy = 0 if copy_orig is None else len(copy_orig)
        self.nn = self.nng + self.ncopy
        self.active = np.asarray(active, bool)
        h = ctypes.c_void_p()
        xe, ye, ze = (np.ascontiguousarray(e, np.float64) for e in (grid.xe, grid.ye, grid.ze))
        with torch.cuda.device(self.device):
            check(self.lib.cm_sys_create(ctypes.byref(h), ctypes.c_void_p(self.stream.cuda_stream), grid.nx, grid.ny,
                                         grid.nz, ptr(xe, ctypes.c_double), ptr(ye, ctypes.c_double),
                                         ptr(ze, ctypes.c_double), int(self.sri)))
        self.h = h
        self.materials = materials
        try:
            self._describe(labels, lam_tab, mu_tab, lam, mu, copy_orig, irr_cells, irr_nodes, extra)
            if rank1 is not None:
                rp, rd, rw, rk = (np.ascontiguousarray(a, t) for a, t in
                                  zip(rank1, (np.int64, np.int64, np.float64, np.float64)))
                check(self.lib.cm_sys_set_rank1(self.h, len(rk), ptr(rp, ctypes.c_longlong), ptr(rd, ctypes.c_longlong),
                                                ptr(rw, ctypes.c_double), ptr(rk, ctypes.c_double)))
            if used is not None:
                um = np.ascontiguousarray(used, np.uint8)
                check(self.lib.cm_sys_set_used(self.h, len(um), ptr(um, ctypes.c_uint8)))
            fixed = 0 if levels is None else int(levels)
            check(self.lib.cm_sys_set_last_coarsening(self.h, int(last_axes)))
            check(self.lib.cm_sys_set_copy_blocks(self.h, int(bool(copy_blocks))))
            check(self.lib.cm_sys_set_march_blocks(self.h, int(march_blocks)))
            with torch.cuda.device(self.device):
                check(self.lib.cm_sys_build(self.h, int(min_dofs), int(max_levels), fixed, SMOOTHERS[smoother],
                                            int(degree), float(omega), int(npre), int(npost)))
        except Exception:
            self.close()
            raise
        self.smoother = smoother

    # ------------------------------------------------------------------ description
    def _describe(self, labels, lam_tab, mu_tab, lam, mu, copy_orig, irr_cells, irr_nodes, extra):
        lib = self.lib
        if labels is None and self.materials != "values":
            comp = compress_materials(lam, mu, self.active)
            if comp is None and self.materials == "labels":
                raise ValueError("more than 255 distinct (lam, mu) pairs: use materials='values'")
            if comp is not None:
                lam_tab, mu_tab, labels = comp
        if labels is not None:
            lt = np.ascontiguousarray(lam_tab, np.float64)
            mt = np.ascontiguousarray(mu_tab, np.float64)
            lb = np.ascontiguousarray(labels, np.uint8).copy()
            lb[~self.active] = 255
            check(lib.cm_sys_set_labels(self.h, len(lt), ptr(lt, ctypes.c_double), ptr(mt, ctypes.c_double),
                                        ptr(lb, ctypes.c_uint8)))
            self.label_mode = True
            self.lam_tab, self.mu_tab, self.labels = lt, mt, lb
        else:
            la = np.ascontiguousarray(lam, np.float64)
            ma = np.ascontiguousarray(mu, np.float64)
            ac = np.ascontiguousarray(self.active, np.uint8)
            check(lib.cm_sys_set_cellmat(self.h, ptr(la, ctypes.c_double), ptr(ma, ctypes.c_double),
                                         ptr(ac, ctypes.c_uint8)))
            self.label_mode = False
        co = np.ascontiguousarray(np.zeros(0) if copy_orig is None else copy_orig, np.int64)
        ic = np.ascontiguousarray(np.zeros(0) if irr_cells is None else irr_cells, np.int64)
        inod = np.ascontiguousarray(np.zeros((0, 8)) if irr_nodes is None else irr_nodes, np.int64)
        check(lib.cm_sys_set_split(self.h, len(co), ptr(co, ctypes.c_longlong), len(ic), ptr(ic, ctypes.c_longlong),
                                   ptr(inod, ctypes.c_longlong)))
        self.irr_cells, self.irr_nodes = ic, inod
        if extra is not None:
            r, c, v = (np.ascontiguousarray(a, t) for a, t in zip(extra, (np.int64, np.int64, np.float64)))
            check(lib.cm_sys_set_extra(self.h, len(r), ptr(r, ctypes.c_longlong), ptr(c, ctypes.c_longlong),
                                       ptr(v, ctypes.c_double)))

    # ------------------------------------------------------------------ info
    def levels(self):
        cap = 16
        desc = np.zeros(6 * cap, np.int64)
        n = ctypes.c_int()
        check(self.lib.cm_sys_levels(self.h, ctypes.byref(n), ptr(desc, ctypes.c_longlong), cap))
        d = desc[: 6 * n.value].reshape(-1, 6)
        return [dict(nx=int(r[0]), ny=int(r[1]), nz=int(r[2]), nodes=int(r[3]), ndof=int(r[4]), extra_rows=int(r[5]))
                for r in d]

    def level_stats(self, level):
        """Operator structure of one level: cell, irregular-cell, extra-row and rank-one counts, the
        coarse per-node list lengths applied inside the structured march, the z extent of a march
        block per precision for products (kz_f64, kz_f32) and fused smoothing steps, and (level 1)
        the cells with several materials, applied by the heterogeneous-cell pass."""
        st = np.zeros(15, np.int64)
        check(self.lib.cm_sys_level_stats(self.h, int(level), ptr(st, ctypes.c_longlong), 15))
        keys = ("cells", "irregular", "extra_rows", "extra_nnz", "rank1_terms", "rank1_nnz", "block_entries",
                "block_row_max", "rank1_node_entries", "rank1_node_max", "kz_f64", "kz_f64_smooth", "kz_f32",
                "kz_f32_smooth", "het_cells")
        return dict(zip(keys, (int(v) for v in st)))

    def timings(self):
        """Build stage times (ms)."""
        t = np.zeros(9)
        check(self.lib.cm_sys_timings(self.h, ptr(t, ctypes.c_double), 9))
        keys = ("prepare", "structure", "upload", "moments", "diag", "coarse_inverse", "total", "structure_used",
                "structure_extra")
        return dict(zip(keys, (round(float(v), 3) for v in t)))

    def zeros(self, dtype=torch.float64, level=0):
        nn = self.nn if level == 0 else self.levels()[level]["nodes"]
        return torch.zeros(3 * nn, dtype=dtype, device=self.device)

    # ------------------------------------------------------------------ compute
    def _ck(self, t, dtype, n):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous() and t.numel() == n):
            raise ValueError(f"expected a contiguous CUDA {dtype} tensor of {n} values")

    def apply(self, u, y=None, level=0):
        prec = PREC[u.dtype]
        n = u.numel()
        y = torch.empty_like(u) if y is None else y
        self._ck(u, u.dtype, n)
        self._ck(y, u.dtype, n)
        check(self.lib.cm_sys_apply(self.h, level, prec, ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(y.data_ptr())))
        return y

    def vcycle(self, b, x=None):
        prec = PREC[b.dtype]
        x = torch.empty_like(b) if x is None else x
        self._ck(b, b.dtype, 3 * self.nn)
        self._ck(x, b.dtype, 3 * self.nn)
        check(self.lib.cm_sys_vcycle(self.h, prec, ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr())))
        return x

    def solve(self, b, x0=None, tol=1e-8, maxiter=800, mode="fp64", check_every=1, inner_tol=1e-3):
        """PCG on A x = b (FP64 device vectors, level-0 layout).  Returns (x, info)."""
        n = 3 * self.nn
        self._ck(b, torch.float64, n)
        x = torch.zeros(n, dtype=torch.float64, device=self.device) if x0 is None else x0.clone()
        self._ck(x, torch.float64, n)
        info = _lib.SolveInfo()
        check(self.lib.cm_sys_solve(self.h, MODES[mode], ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                    float(tol), int(maxiter), int(check_every), float(inner_tol), ctypes.byref(info)))
        out = dict(iters=info.iters, flag=info.flag, outer=info.outer, levels=info.levels, relres=info.relres,
                   t_solve_ms=info.t_solve_ms, t_setup_ms=info.t_setup_ms, mode=mode, device=str(self.device))
        return x, out

    def prepare_solve(self):
        """Everything a mode-"ref" solve needs before its first launch (work vectors, FP32 V-cycle
        data, FP64 operator data, block orders), done now on this system's stream; solves and
        lock-step batches then start without a wait on the host."""
        check(self.lib.cm_sys_prepare_solve(self.h))

    def bounds(self, level, dtype=torch.float64):
        """Chebyshev bounds (lmax, lmin) of one level (fe/gmg.py:57-58)."""
        out = np.zeros(2)
        check(self.lib.cm_sys_level_data(self.h, int(level), PREC[dtype], 0, out.ctypes.data_as(ctypes.c_void_p)))
        return float(out[0]), float(out[1])

    def block_inverses(self, level, dtype=torch.float64):
        """Block-Jacobi inverses of one level as (nodes, 3, 3) (fe/gmg.py:29-44)."""
        nn = self.nn if level == 0 else self.levels()[level]["nodes"]
        d = torch.empty(6 * nn, dtype=dtype, device=self.device)
        check(self.lib.cm_sys_level_data(self.h, int(level), PREC[dtype], 1, ctypes.c_void_p(d.data_ptr())))
        v = d.view(nn, 6).double().cpu().numpy()
        idx = np.array([[0, 1, 2], [1, 3, 4], [2, 4, 5]])
        return v[:, idx]

    def power(self, level, x0):
        """12 power steps of Dinv A from x0 (device vector, the level's layout): sets the level's
        Chebyshev bounds to 1.1 lam and lmax / 30 and returns lam (fe/gmg.py:51-58)."""
        nn = self.nn if level == 0 else self.levels()[level]["nodes"]
        self._ck(x0, x0.dtype, 3 * nn)
        lam = ctypes.c_double()
        check(self.lib.cm_sys_power(self.h, int(level), PREC[x0.dtype], ctypes.c_void_p(x0.data_ptr()),
                                    ctypes.byref(lam)))
        return lam.value

    def reference_bounds(self, level0_nodes=None, dtype=torch.float64):
        """Chebyshev bounds of every smoothed level from the reference's start vector: the Level
        constructor of fe/gmg.py:51 draws numpy.random.default_rng(0).standard_normal(n) over the
        level's compact DOFs (level 0: node ``level0_nodes[i]`` holds compact DOFs 3i..3i+2, default
        the used nodes in ascending order; coarser levels: their used nodes in ascending order).
        Returns the estimates lam per level."""
        lv = self.levels()
        lams = []
        for l in range(len(lv) - 1):
            nn = self.nn if l == 0 else lv[l]["nodes"]
            nodes = level0_nodes if (l == 0 and level0_nodes is not None) else np.flatnonzero(self.used_mask(l))
            nodes = np.asarray(nodes, np.int64)
            v = np.random.default_rng(0).standard_normal(3 * len(nodes))
            x = np.zeros(3 * nn)
            x[(3 * nodes[:, None] + np.arange(3)).ravel()] = v
            lams.append(self.power(l, torch.tensor(x, dtype=dtype, device=self.device)))
        return lams

    def used_mask(self, level):
        """Nodes of one level that carry unknowns (the reference's kept DOFs / 3)."""
        nn = self.nn if level == 0 else self.levels()[level]["nodes"]
        m = np.zeros(nn, np.uint8)
        check(self.lib.cm_sys_level_data(self.h, int(level), 0, 2, m.ctypes.data_as(ctypes.c_void_p)))
        return m.astype(bool)

    def smooth(self, b, x=None, level=0, zero_start=False):
        """One call of the level's smoother in place (fe/gmg.py:61-76): x = smooth(x, b); with
        zero_start, x = smooth(0, b).  Returns x."""
        nn = self.nn if level == 0 else self.levels()[level]["nodes"]
        self._ck(b, b.dtype, 3 * nn)
        if x is None:
            x = torch.zeros_like(b)
            zero_start = True
        self._ck(x, b.dtype, 3 * nn)
        check(self.lib.cm_sys_smooth(self.h, int(level), PREC[b.dtype], ctypes.c_void_p(b.data_ptr()),
                                     ctypes.c_void_p(x.data_ptr()), int(bool(zero_start))))
        return x

    def time_apply(self, u, y, nrep=20, structured_only=False, level=0, smooth_step=False):
        """Average ms of nrep launches (HIP events): the whole product, the structured stencil kernel
        alone (structured_only), or the structured kernel of one fused Chebyshev step (smooth_step)."""
        ms = ctypes.c_double()
        what = 2 if smooth_step else int(bool(structured_only))
        check(self.lib.cm_sys_time_apply(self.h, level, PREC[u.dtype], ctypes.c_void_p(u.data_ptr()),
                                         ctypes.c_void_p(y.data_ptr()), int(nrep), what, ctypes.byref(ms)))
        return ms.value

    def autotune(self, precisions=("f32", "f64"), nrep=5, retime=False):
        """Pick the fastest structured-block z extent per level (project autotuner, HIP events);
        retime: time level shapes the process has tuned before again."""
        for p in precisions:
            check(self.lib.cm_sys_autotune(self.h, PREC[p], -int(nrep) if retime else int(nrep)))

    def tuning(self):
        """z extents of every level's march blocks (the autotuner's or the heuristic's):
        [[nx, ny, nz, kz_f64, kz_f64_smooth, kz_f32, kz_f32_smooth], ...] for the levels with a
        smoother."""
        out = []
        for l, lv in enumerate(self.levels()[:-1]):
            st = self.level_stats(l)
            out.append([lv["nx"], lv["ny"], lv["nz"]] + [st[k] for k in ("kz_f64", "kz_f64_smooth", "kz_f32",
                                                                        "kz_f32_smooth")])
        return out

    def set_tuning(self, table):
        """Apply a table of ``tuning()``'s form to the levels whose shape matches (others keep theirs)."""
        lvs = self.levels()
        for l, row in enumerate(table):
            if l >= len(lvs) - 1 or [lvs[l]["nx"], lvs[l]["ny"], lvs[l]["nz"]] != list(row[:3]):
                continue
            for prec, smooth, kz in ((0, 0, row[3]), (0, 1, row[4]), (1, 0, row[5]), (1, 1, row[6])):
                check(self.lib.cm_sys_set_tuned_kz(self.h, l, prec, smooth, int(kz)))

    def set_kz(self, kz):
        check(self.lib.cm_sys_set_kz(self.h, int(kz)))

    def set_block_skip(self, on=True):
        """Structured-march blocks without stiffness run only the epilogue (default on)."""
        check(self.lib.cm_sys_set_block_skip(self.h, int(bool(on))))

    def set_graphs(self, on=True):
        """Replay the V-cycle through a captured HIP graph (default) or launch it eagerly."""
        check(self.lib.cm_sys_set_graphs(self.h, int(bool(on))))

    def set_rigid(self, rows, cols, vals, ncol, ordered=False):
        """Rigid-mode correction Z Einv Z^T (fe/gmg.py:127-135): Z as COO over level-0 DOFs
        (ordered: already in (row, column) order).  E = sym(Z^T A Z) is formed with the GPU
        operator, Einv = pinv(E) (ncol x ncol, host)."""
        rows = np.asarray(rows, np.int64)
        cols = np.asarray(cols, np.int64)
        vals = np.asarray(vals, np.float64)
        ncol = int(ncol)
        if ncol == 0 or len(rows) == 0:
            return
        if not ordered:   # (row, column) order: the library then skips its sort
            rc = np.lexsort((cols, rows))
            rows, cols, vals = (a[rc] for a in (rows, cols, vals))
        rows, cols, vals = (np.ascontiguousarray(a) for a in (rows, cols, vals))
        n = 3 * self.nn
        # E = Z^T A Z one column at a time: scatter z_j into one scratch vector, apply A, gather
        # Z^T (A z_j) from the sparse entries (Z holds a few entries per node and column)
        order = np.argsort(cols, kind="stable")
        rows_s, cols_s, vals_s = rows[order], cols[order], vals[order]
        bounds = np.searchsorted(cols_s, np.arange(ncol + 1))
        rd = torch.from_numpy(rows_s).to(self.device)
        vd = torch.from_numpy(vals_s).to(self.device)
        cd = torch.from_numpy(cols_s).to(self.device)
        z = torch.zeros(n, dtype=torch.float64, device=self.device)
        az = torch.empty_like(z)
        Ed = torch.zeros((ncol, ncol), dtype=torch.float64, device=self.device)
        for j in range(ncol):
            a, b = int(bounds[j]), int(bounds[j + 1])
            if a == b:
                continue
            z.zero_()
            z.index_put_((rd[a:b],), vd[a:b], accumulate=True)
            self.apply(z, az)
            Ed[:, j].index_add_(0, cd, vd * az[rd])
        E = Ed.cpu().numpy()
        Einv = np.ascontiguousarray(np.linalg.pinv(0.5 * (E + E.T)))
        check(self.lib.cm_sys_set_rigid(self.h, int(ncol), len(rows), ptr(rows, ctypes.c_longlong),
                                        ptr(cols, ctypes.c_longlong), ptr(vals, ctypes.c_double),
                                        ptr(Einv, ctypes.c_double)))
        self.rigid_cols = ncol

    def clear_rigid(self):
        """Remove the additive coarse correction set by set_rigid."""
        check(self.lib.cm_sys_set_rigid(self.h, 0, 0, None, None, None, None))
        self.rigid_cols = 0

    def set_fallback(self, iters=5, drop=10.0):
        """Early exit of a solve from a predicted x0: unless the FP64 true residual has dropped by
        ``drop`` after ``iters`` iterations, solve() stops with flag -2 (iters = 0: off)."""
        check(self.lib.cm_sys_set_fallback(self.h, int(iters), float(drop)))

    def sync(self):
        check(self.lib.cm_sys_sync(self.h))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.cm_sys_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cell_stress_gpu(grid, lam, mu, U, cells, nodes8=None, device=None, coeffs=None):
    """Cell-centre stress/strain/von Mises on the GPU (fe/solver.py:257-272).  U: host (nnode,3)
    or device tensor; coeffs: (ncell, 6) z-orthotropic coefficients instead of lam/mu; returns host
    arrays (sig, vm, eps)."""
    dev = require_gpu(device)
    lib = _lib.load()
    cells = np.asarray(cells, np.int64)
    n = len(cells)
    if n == 0:
        return np.zeros((0, 6)), np.zeros(0), np.zeros((0, 6))
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    Ud = U if isinstance(U, torch.Tensor) else t(np.asarray(U, float).reshape(-1), torch.float64)
    hx, hy, hz = t(grid.dx, torch.float64), t(grid.dy, torch.float64), t(grid.dz, torch.float64)
    cd = t(cells, torch.int64)
    nd = None if nodes8 is None else t(np.asarray(nodes8, np.int64).reshape(n, 8), torch.int64)
    la, ma = t(np.asarray(lam, float), torch.float64), t(np.asarray(mu, float), torch.float64)
    sig = torch.empty(n * 6, dtype=torch.float64, device=dev)
    vm = torch.empty(n, dtype=torch.float64, device=dev)
    eps = torch.empty(n * 6, dtype=torch.float64, device=dev)
    vp = lambda x: ctypes.c_void_p(0 if x is None else x.data_ptr())
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    if coeffs is not None:
        cf = t(np.asarray(coeffs, float).reshape(-1, 6), torch.float64)
        check(lib.cm_cell_stress_ortho(st, grid.nx, grid.ny, grid.nz, vp(hx), vp(hy), vp(hz), n, vp(cd), vp(nd), vp(cf),
                                       vp(Ud), vp(sig), vp(vm), vp(eps)))
    else:
        check(lib.cm_cell_stress(st, grid.nx, grid.ny, grid.nz, vp(hx), vp(hy), vp(hz), n, vp(cd), vp(nd), vp(la),
                                 vp(ma), vp(Ud), vp(sig), vp(vm), vp(eps)))
    return sig.view(n, 6).cpu().numpy(), vm.cpu().numpy(), eps.view(n, 6).cpu().numpy()
