"""Drop-in for ``spring_matrix_cn`` of the reference's ``rb/replica3.py:68-86`` (ligament springs).

Returns the reference's callback ``build(nmap, n) -> CSR`` on compact DOFs: per spring with unit
direction d = (b - a)/|b - a|, w = +d/n_A on the n_A end-A patch nodes and -d/n_B on the end-B
patch (patch = the bone's used nodes within max(4 mm, distance of the 8th nearest)), K += k w w^T.
``solve_cn`` turns it into sparse extra stiffness applied on the GPU.  ``spring_pairs`` is the
single-node form (w = +d on node a, -d on node b) used by the synthetic configurations.
"""
import numpy as np
import scipy.sparse as sp


def spring_matrix_cn(xyz, springs, bone_nodes):
    def build(nmap, n):
        rows, cols, vals = [], [], []
        for s in springs:
            a = np.asarray(s["a"], float)
            b = np.asarray(s["b"], float)
            d = (b - a) / np.linalg.norm(b - a)
            dofs, w = [], []
            for p, sign, key in ((a, 1.0, "A"), (b, -1.0, "B")):
                cand = np.asarray(bone_nodes[s[key]], np.int64)
                cand = cand[nmap[cand] >= 0]
                dist = np.linalg.norm(xyz[cand] - p, axis=1)
                radius = max(4.0, np.sort(dist)[min(7, len(dist) - 1)])
                ln = nmap[cand[dist <= radius]]
                for c in range(3):
                    dofs.append(3 * ln + c)
                    w.append(np.full(len(ln), sign * d[c] / len(ln)))
            dofs = np.concatenate(dofs)
            w = np.concatenate(w)
            ud, inv = np.unique(dofs, return_inverse=True)
            wu = np.zeros(len(ud))
            np.add.at(wu, inv, w)
            keep = wu != 0
            ud, wu = ud[keep], wu[keep]
            rows.append(np.repeat(ud, len(ud))); cols.append(np.tile(ud, len(ud)))
            vals.append(s["k"] * np.outer(wu, wu).ravel())
        if not rows:
            return sp.csr_matrix((n, n))
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return build


def spring_pairs(xyz, pairs, k, directional=True):
    """Springs between single nodes as COO over node DOFs 3*node+c.  directional: k (d.(u_a-u_b))^2
    with d the unit vector a->b (ligament form); otherwise isotropic k |u_a - u_b|^2 (interface form)."""
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    k = np.broadcast_to(np.asarray(k, float), (len(pairs),))
    rows, cols, vals = [], [], []
    for (a, b), kk in zip(pairs, k):
        if directional:
            d = xyz[b] - xyz[a]
            d = d / np.linalg.norm(d)
            dofs = np.r_[3 * a + np.arange(3), 3 * b + np.arange(3)]
            w = np.r_[d, -d]
            rows.append(np.repeat(dofs, 6)); cols.append(np.tile(dofs, 6)); vals.append(kk * np.outer(w, w).ravel())
        else:
            for c in range(3):
                da, db = 3 * a + c, 3 * b + c
                rows.append([da, db, da, db]); cols.append([da, db, db, da]); vals.append([kk, kk, -kk, -kk])
    if not rows:
        z = np.zeros(0, np.int64)
        return z, z, np.zeros(0)
    return (np.concatenate(rows).astype(np.int64), np.concatenate(cols).astype(np.int64),
            np.concatenate(vals).astype(float))


def spring_terms(xyz, pairs, k, directional=True):
    """Springs between single nodes as rank-one terms (ptr, dofs, w, k) over node DOFs: directional
    -> one term w = (+d at a, -d at b); isotropic -> one term per component, w = e_a - e_b."""
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    kk = np.broadcast_to(np.asarray(k, float), (len(pairs),))
    a, b = pairs[:, 0], pairs[:, 1]
    if directional:
        d = xyz[b] - xyz[a]
        d = d / np.linalg.norm(d, axis=1, keepdims=True)
        dofs = np.concatenate([3 * a[:, None] + np.arange(3), 3 * b[:, None] + np.arange(3)], 1)
        w = np.concatenate([d, -d], 1)
        ptr = np.arange(len(pairs) + 1) * 6
        return ptr.astype(np.int64), dofs.ravel().astype(np.int64), w.ravel(), kk.copy()
    dofs = np.stack([3 * a[:, None] + np.arange(3), 3 * b[:, None] + np.arange(3)], 2)   # (m, 3, 2)
    w = np.broadcast_to(np.array([1.0, -1.0]), dofs.shape)
    ptr = np.arange(3 * len(pairs) + 1) * 2
    return ptr.astype(np.int64), dofs.reshape(-1).astype(np.int64), w.reshape(-1).copy(), np.repeat(kk, 3)


def cat_terms(*terms):
    """Concatenate rank-one term sets (ptr, dofs, w, k)."""
    terms = [t for t in terms if t is not None and len(t[3])]
    if not terms:
        return None
    ptr, off = [np.zeros(1, np.int64)], 0
    for t in terms:
        ptr.append(t[0][1:] + off)
        off += t[0][-1]
    return (np.concatenate(ptr), np.concatenate([t[1] for t in terms]), np.concatenate([t[2] for t in terms]),
            np.concatenate([t[3] for t in terms]))
