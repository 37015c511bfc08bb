"""Inputs and targets of the learned warm start, built on the GPU from the solver's level-0
description (``synth.level0_description``: labels and (lambda, mu) tables, split copies and the
cells touching them, ground springs, spring terms, nodal loads).

Input channels per cell (``cm_nn_features``): active, a one-hot material class (E about 1, 5, 700
and 17000 MPa: soft tissue, filled gap, cancellous/coalition, cortical), (nu - 0.3) / 0.15, the
interface-spring stiffness on cells touching split copies, ground-spring and ligament-end weights
and the downward load on the cell's corners, and the normalised cell centre.  The target is the
cell-mean displacement divided by the total load (the system is linear in the load) times a fixed
scale, so that load-scaled actions share one target.
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from .._lib import check_nn
from .net import IN_CH


def _vp(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def material_classes(lam_tab, mu_tab):
    """(class per label (int8, 256 entries, -1 unused), nu per label (float32, 256))."""
    lam, mu = np.asarray(lam_tab, float), np.asarray(mu_tab, float)
    E = mu * (3 * lam + 2 * mu) / np.maximum(lam + mu, 1e-300)
    nu = lam / np.maximum(2 * (lam + mu), 1e-300)
    cls = np.full(256, -1, np.int8)
    cls[:len(E)] = np.select([E < 3.0, E < 50.0, E < 5000.0], [0, 1, 2], 3).astype(np.int8)
    nus = np.full(256, 0.3, np.float32)
    nus[:len(nu)] = nu
    return cls, nus


def encode(desc, device):
    """Device inputs of one system: features F (BF16 as int16 [V][16]) and what the cell <-> node
    maps need (labels, irregular-cell index, node lists), the grid sizes and the total load."""
    dev = torch.device(device)
    lib = _lib.load()
    nx, ny, nz = (len(e) - 1 for e in desc["edges"])
    V, nnode = nx * ny * nz, (nx + 1) * (ny + 1) * (nz + 1)
    copy_orig = np.asarray(desc["copy_orig"], np.int64)

    def to_grid(node):                 # a split copy counts at its original grid node
        node = np.asarray(node, np.int64)
        if len(copy_orig) == 0:
            return node
        return np.where(node >= nnode, copy_orig[np.clip(node - nnode, 0, len(copy_orig) - 1)], node)
    cls, nus = material_classes(desc["lam_tab"], desc["mu_tab"])
    split = np.zeros(V, np.uint8)
    split[np.asarray(desc["irr_cells"], np.int64)] = 1
    ground = np.zeros(nnode, np.float32)
    if desc["extra"] is not None and len(desc["extra"][0]):
        r, c, v = (np.asarray(a) for a in desc["extra"])
        diag = r == c
        nodes = to_grid(r[diag] // 3)
        np.maximum.at(ground, nodes, (np.log10(1.0 + np.abs(v[diag])) / 6.0).astype(np.float32))
    lig = np.zeros(nnode, np.float32)
    kif = 0.0
    if desc["terms"] is not None and len(desc["terms"][3]):
        ptr, dofs, w, k = (np.asarray(a) for a in desc["terms"])
        term_of = np.repeat(np.arange(len(k)), np.diff(ptr))
        node = dofs // 3
        touches_copy = np.zeros(len(k), bool)
        np.logical_or.at(touches_copy, term_of, node >= nnode)
        if touches_copy.any():
            kif = float(np.log10(np.max(k[touches_copy])) / 6.0)
        lig_terms = ~touches_copy
        lig[to_grid(node[lig_terms[term_of]])] = 1.0
    load = np.zeros(nnode, np.float32)
    fi, fv = np.asarray(desc["f_idx"], np.int64), np.asarray(desc["f_val"], float)
    zsel = fi % 3 == 2
    np.add.at(load, to_grid(fi[zsel] // 3), fv[zsel].astype(np.float32))
    ftot = float(np.abs(fv).sum())
    lmax = float(np.abs(load).max()) if len(load) else 0.0
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    # every device input stays referenced by the returned dict, so no buffer goes back to the
    # allocator while a kernel that reads it may still be queued
    keep = dict(lab=t(desc["labels"], torch.uint8), cls=t(cls, torch.int8), nu=t(nus, torch.float32),
                split=t(split, torch.uint8), ground=t(ground, torch.float32), lig=t(lig, torch.float32),
                load=t(load, torch.float32))
    F = torch.empty(V * IN_CH, dtype=torch.int16, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    check_nn(lib.cm_nn_features(st, nx, ny, nz, _vp(keep["lab"]), _vp(keep["cls"]), _vp(keep["nu"]),
                                _vp(keep["split"]), kif, _vp(keep["ground"]), _vp(keep["lig"]), _vp(keep["load"]),
                                1.0 / lmax if lmax > 0 else 0.0, _vp(F)))
    irr_of = np.full(V, -1, np.int32)
    irr_of[np.asarray(desc["irr_cells"], np.int64)] = np.arange(len(desc["irr_cells"]), dtype=np.int32)
    irr_nodes = np.asarray(desc["irr_nodes"], np.int64).reshape(-1, 8)
    if len(irr_nodes) == 0:
        irr_nodes = np.zeros((1, 8), np.int64)
    lab = keep["lab"]
    if "part" in desc:
        keep["part"] = t(desc["part"], torch.uint8)
    keep["copy_orig"] = t(copy_orig if len(copy_orig) else np.zeros(1, np.int64), torch.int64)
    keep["edges"] = [t(np.asarray(e, float), torch.float64) for e in desc["edges"]]
    return dict(F=F, lab=lab, irr_of=t(irr_of, torch.int32), irr_nodes=t(irr_nodes, torch.int64), dims=(nx, ny, nz),
                nn=int(desc["nn"]), ftot=ftot, mask=(lab != 255).to(torch.uint8), inputs=keep, kif=kif,
                load_scale=1.0 / lmax if lmax > 0 else 0.0)


def targets(enc, x, tscale):
    """Cell targets (FP32 [V][3]): the cell-mean of the solution x (FP64 node vector) x tscale /
    total load."""
    lib = _lib.load()
    nx, ny, nz = enc["dims"]
    t = torch.empty(nx * ny * nz * 3, dtype=torch.float32, device=x.device)
    st = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    check_nn(lib.cm_nn_node_to_cell(st, nx, ny, nz, _vp(enc["lab"]), _vp(enc["irr_of"]), _vp(enc["irr_nodes"]),
                                    _vp(x), float(tscale / max(enc["ftot"], 1e-300)), _vp(t)))
    return t


def to_nodes(enc, pred, tscale):
    """FP64 node vector (3 per node, the solver's level-0 layout) from the network's per-cell output
    (FP32 [V][16], first 3 channels): each node takes the mean over its active cells, rescaled by
    the total load / tscale; unused nodes 0."""
    lib = _lib.load()
    nx, ny, nz = enc["dims"]
    dev = pred.device
    nn = enc["nn"]
    acc = enc.get("acc")
    if acc is None or acc.numel() != 4 * nn or acc.device != dev:
        acc = enc["acc"] = torch.empty(4 * nn, dtype=torch.float32, device=dev)
    x = torch.empty(3 * nn, dtype=torch.float64, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    check_nn(lib.cm_nn_cell_to_node(st, nx, ny, nz, _vp(enc["lab"]), _vp(enc["irr_of"]), _vp(enc["irr_nodes"]),
                                    _vp(pred), 16, nn, _vp(acc), float(enc["ftot"] / tscale), _vp(x)))
    return x
