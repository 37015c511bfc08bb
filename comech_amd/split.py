"""Split-node numbering on the host (numpy only): the node ids of rb/splitfe.py:261-305.

``comech_amd.rb.splitfe`` re-exports ``split_connectivity_fast`` under the reference's module name;
the GPU kernel ``cm_split_connectivity`` computes the same numbering on the device.
"""
import numpy as np

from .grid import corners_of


def split_connectivity_fast(g, active, part, bone_share):
    """Split connectivity with the numbering of rb/splitfe.py:261-305 (the loop form :17-55 gives
    the same result).

    Rule: a grid node touched by active cells of two or more bones is split.  Its lowest bone keeps
    the grid id; every further (node, bone) pair gets ``nnode + rank``, ranked in (node, bone)
    order.  A bone cell uses the id of its own bone at each corner.  A part-0 (joint/soft) cell at a
    split node uses the node's bone with the largest share in that cell (the first in bone order
    on a tie), or, when none of the node's bones has a positive share there, the node's bone with
    the most cells at the node (again the first on a tie).

    Computed from a dense table count[node, bone] over the nodes bone cells touch: the split
    nodes are its rows with two or more non-zero columns, the copy ids a running count over the
    row-major table, and the soft-cell choice two masked argmax over that row.
    Returns (cn (ncell,8), orig (nnode+ncopy,), number of split nodes)."""
    cells = np.flatnonzero(active)
    corn = corners_of(g, cells)
    pc = np.asarray(part)[cells].astype(np.int64)
    nb = int(bone_share.shape[1])
    cn = g.cell_nodes.copy()
    bc = pc > 0
    if not bc.any():
        return cn, np.arange(g.nnode, dtype=np.int64), 0
    bnodes = corn[bc]
    touched = np.zeros(g.nnode, bool)
    touched[bnodes.ravel()] = True
    tn = np.flatnonzero(touched)
    row_of = np.full(g.nnode, -1, np.int64)
    row_of[tn] = np.arange(len(tn))
    count = np.bincount((row_of[bnodes] * nb + pc[bc][:, None]).ravel(),
                        minlength=len(tn) * nb).reshape(len(tn), nb)
    count[:, 0] = 0
    rows = np.flatnonzero((count > 0).sum(1) >= 2)
    split_nodes = tn[rows]
    cnt = count[rows]                                   # (nsplit, nb)
    has = cnt > 0
    lowest = has.argmax(1)
    ids = np.full(has.shape, -1, np.int64)
    more = has.copy()
    more[np.arange(len(rows)), lowest] = False
    flat = np.flatnonzero(more.ravel())                 # row-major: (node, bone) order
    ids.ravel()[flat] = g.nnode + np.arange(len(flat))
    ids[np.arange(len(rows)), lowest] = split_nodes
    orig = np.concatenate([np.arange(g.nnode, dtype=np.int64), np.repeat(split_nodes, more.sum(1))])
    srow = np.full(g.nnode, -1, np.int64)
    srow[split_nodes] = np.arange(len(rows))
    q = srow[corn]
    r, l = np.nonzero(q >= 0)
    if len(r):
        s_ = q[r, l]
        bone = pc[r].copy()
        soft = np.flatnonzero(bone == 0)
        if len(soft):
            cand = has[s_[soft]]
            share = np.where(cand, np.asarray(bone_share)[cells[r[soft]]], -np.inf)
            by_share = share.argmax(1)
            top = share[np.arange(len(soft)), by_share]
            by_count = np.where(cand, cnt[s_[soft]], -1).argmax(1)
            bone[soft] = np.where(top > 0, by_share, by_count)
        cn[cells[r], l] = ids[s_, bone]
    return cn, orig, len(rows)


def irregular_cells(grid, active, cn):
    """Active cells whose connectivity differs from the grid corners (cells touching a split copy).
    Returns (cells (n,), their 8 node ids (n, 8))."""
    cells = np.flatnonzero(active)
    out_c, out_n = [], []
    for a in range(0, len(cells), 1 << 20):
        c = cells[a:a + (1 << 20)]
        row = cn[c]
        diff = (row != corners_of(grid, c)).any(1)
        out_c.append(c[diff]); out_n.append(row[diff])
    if not out_c:
        return np.zeros(0, np.int64), np.zeros((0, 8), np.int64)
    return np.concatenate(out_c).astype(np.int64), np.concatenate(out_n).astype(np.int64)
