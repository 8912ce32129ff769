"""Element tables and element-pair lists: the data the operator plan consumes.

Host-side builders for the reference's x/y row order and connectivity
(project-asgard/DG-SparseGrid): the new hash (hash_table_sparse_nD.m, default),
the old hash (hash_table_nD.m, needed for connectivity), global_matrix.m's
per-element 1D indices and connect_1D.m / connect_nD.m's maximal pair lists.
The enumerations follow the reference's loop orders; the pair lists are built
with a vectorised membership test instead of the reference's per-tuple hash
lookups (same sorted sets).
"""
from dataclasses import dataclass

import numpy as np


def lev_cell_to_1d(lev, cell):
    """lev_cell_to_1D_index.m:6-9 -- 1-based hierarchical 1D index (vectorised)."""
    lev = np.asarray(lev, dtype=np.int64)
    cell = np.asarray(cell, dtype=np.int64)
    idx = np.where(lev == 0, 1, (np.int64(1) << np.maximum(lev - 1, 0)) + cell + 1)
    return idx.astype(np.int64)


# ------------------------------------------------------------- enumerations
def perm_eq_d(idim, levmax, sumeq, last_index_decreasing=False):
    """PERMUTE/perm_eq_d.m: tuples with sum == sumeq, entry i <= levmax[i]."""
    levmax = list(np.broadcast_to(np.asarray(levmax), (idim,)))
    if idim == 1:
        return [(sumeq,)] if levmax[0] >= sumeq else []
    top = min(levmax[-1], sumeq)
    order = range(top, -1, -1) if last_index_decreasing else range(0, top + 1)
    out = []
    for last in order:
        out += [t + (last,) for t in perm_eq_d(idim - 1, levmax[:-1], sumeq - last, last_index_decreasing)]
    return out


def perm_leq_d(idim, levmax, summax, increasing_sum_order=False):
    """PERMUTE/perm_leq_d.m: all sums, SumMax down to 0 unless increasing."""
    sums = range(0, summax + 1) if increasing_sum_order else range(summax, -1, -1)
    out = []
    for s in sums:
        out += perm_eq_d(idim, levmax, s)
    return out


def perm_eq(idim, n, last_index_decreasing=False):
    """PERMUTE/perm_eq.m."""
    if idim == 1:
        return [(n,)]
    out = []
    for i in range(n + 1):
        if last_index_decreasing:
            out += [t + (n - i,) for t in perm_eq(idim - 1, i, True)]
        else:
            out += [t + (i,) for t in perm_eq(idim - 1, n - i, False)]
    return out


def perm_leq(idim, n, order_by_n=False):
    """PERMUTE/perm_leq.m (old hash order)."""
    if order_by_n:
        out = []
        for i in range(n + 1):
            out += perm_eq(idim, i)
        return out
    if idim == 1:
        return [(i,) for i in range(n + 1)]
    out = []
    for i in range(n + 1):
        out += [t + (i,) for t in perm_leq(idim - 1, n - i)]
    return out


def perm_max(idim, levmax, last_index_decreasing=False):
    """PERMUTE/perm_max.m (full grid)."""
    lv = np.asarray(levmax).ravel()
    lv = [int(lv.max())] * idim if lv.size == 1 else [int(v) for v in lv]
    n = lv[idim - 1]
    rng = range(n, -1, -1) if last_index_decreasing else range(n + 1)
    if idim == 1:
        return [(i,) for i in rng]
    inner = perm_max(idim - 1, lv[: idim - 1], last_index_decreasing)
    return [t + (i,) for i in rng for t in inner]


def _cells(levels):
    """lev_cell_to_1D_index_set.m: all cell tuples of a level tuple, dim 1 fastest."""
    sizes = [1 << max(0, int(l) - 1) for l in levels]
    grids = np.indices(sizes[::-1]).reshape(len(sizes), -1)[::-1]  # last dim slowest
    return grids.T


@dataclass
class ElementTable:
    lev: np.ndarray    # N x d levels
    cell: np.ndarray   # N x d cells
    idx1d: np.ndarray  # N x d 1-based 1D indices (A_data.element_local_index_D)
    ids: np.ndarray    # N md_lev_pos_to_idx ids (1-based; new hash only, else zeros)

    @property
    def num_elem(self):
        return self.lev.shape[0]

    @property
    def ndim(self):
        return self.lev.shape[1]


def _table_from_ptable(ptable, nd, max_lev=None):
    levs, cells = [], []
    for levels in ptable:
        c = _cells(levels)
        levs.append(np.broadcast_to(np.asarray(levels, dtype=np.int64), c.shape))
        cells.append(c)
    lev = np.concatenate(levs).astype(np.int64) if levs else np.zeros((0, nd), np.int64)
    cell = np.concatenate(cells).astype(np.int64) if cells else np.zeros((0, nd), np.int64)
    idx = lev_cell_to_1d(lev, cell)
    ids = np.zeros(lev.shape[0], dtype=np.uint64)
    if max_lev is not None:
        # md_lev_pos_to_idx.m:1-28
        if np.any(lev > max_lev):
            raise ValueError("level exceeds max_lev")
        stride = np.uint64(1) << np.uint64(max_lev)
        acc = np.ones(lev.shape[0], dtype=np.uint64)
        mul = np.uint64(1)
        for d in range(nd):
            acc += (idx[:, d].astype(np.uint64) - np.uint64(1)) * mul
            mul *= stride
        ids = acc
    return ElementTable(lev, cell, idx, ids)


def hash_table_sparse_nd(lev_vec, max_lev, grid_type="SG"):
    """hash_table_sparse_nD.m:1-118 (new hash; descending level-sum order)."""
    lev_vec = [int(v) for v in np.ravel(lev_vec)]
    nd = len(lev_vec)
    ptable = perm_leq_d(nd, lev_vec, max(lev_vec)) if grid_type == "SG" else perm_max(nd, lev_vec)
    return _table_from_ptable(ptable, nd, max_lev)


def hash_table_nd(lev_vec, grid_type="SG"):
    """hash_table_nD.m:1-142 (old hash; uniform level only, :27-29)."""
    lev_vec = [int(v) for v in np.ravel(lev_vec)]
    if len(set(lev_vec)) != 1:
        raise ValueError("use_oldhash not supported for dimension dependent level")
    nd, lev = len(lev_vec), lev_vec[0]
    ptable = perm_leq(nd, lev) if grid_type == "SG" else perm_max(nd, [lev] * nd)
    return _table_from_ptable(ptable, nd, None)


def global_matrix(table):
    """global_matrix.m:44-70 -- A_data: element_local_index_D as an N x d array."""
    return {"element_global_row_index": np.arange(1, table.num_elem + 1),
            "element_local_index_D": table.idx1d.copy()}


# ------------------------------------------------------------ connectivity
def connect_1d(Lev):
    """connect_1D.m:1-51 -- dense 0/1 matrix over 1-based 1D indices (stored 0-based)."""
    n = 1 << Lev
    con = np.zeros((n, n), dtype=bool)

    def link(I, Js):
        Js = np.atleast_1d(Js) - 1
        con[I - 1, Js] = True
        con[Js, I - 1] = True

    for Lx in range(Lev + 1):
        ncx = 1 << max(0, Lx - 1)
        for Px in range(ncx):
            I = int(lev_cell_to_1d(Lx, Px))
            link(I, lev_cell_to_1d(Lx, np.arange(max(Px - 1, 0), min(Px + 1, ncx - 1) + 1)))
            if Px == 0:
                link(I, lev_cell_to_1d(Lx, ncx - 1))
            elif Px == ncx - 1:
                link(I, lev_cell_to_1d(Lx, 0))
            for Ly in range(Lx + 1, Lev + 1):
                nL = Ly - Lx
                ncy = 1 << (Ly - 1)
                w = (1 << nL) if Lx > 0 else (1 << (nL - 1))
                lo, hi = max(w * Px - 1, 0), min(w * Px + w, ncy - 1)
                Py = np.concatenate([[0], np.arange(lo, hi + 1), [ncy - 1]])
                link(I, lev_cell_to_1d(Ly, Py))
    return con


def connect_nd(table, levsum, levmax, grid_type="SG", chunk=4096):
    """connect_nD.m:1-129 over an element table: for every element i the sorted
    list of elements j whose 1D index in every dim is connected to i's
    (connect_1D) and whose levels pass index_leq_max's filter (sum <= levsum,
    max <= levmax; max only for 'FG', index_max.m).

    Returns (row_ptr int64 [N+1], cols int32 1-based)."""
    nd = table.ndim
    Lev = max(levsum, levmax)
    c1 = connect_1d(Lev)
    idx0 = table.idx1d - 1
    if np.any(idx0 >= c1.shape[0]):
        raise ValueError("element 1D index beyond connect_1D(max(levsum, levmax))")
    ok_j = np.all(table.lev <= levmax, axis=1)
    if grid_type == "SG":
        ok_j &= table.lev.sum(axis=1) <= levsum
    N = table.num_elem
    counts = np.zeros(N, dtype=np.int64)
    parts = []
    for r0 in range(0, N, chunk):
        r1 = min(N, r0 + chunk)
        m = np.broadcast_to(ok_j, (r1 - r0, N)).copy()
        for d in range(nd):
            m &= c1[idx0[r0:r1, d]][:, idx0[:, d]]
        rr, cc = np.nonzero(m)  # row-major: per row ascending j
        counts[r0:r1] = np.bincount(rr, minlength=r1 - r0)
        parts.append((cc + 1).astype(np.int32))
    row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cols = np.concatenate(parts) if parts else np.zeros(0, np.int32)
    return row_ptr, cols


# ------------------------------------------------------------ adaptivity
def get_child_elements(lev_vec, pos_vec, max_lev, refinement_method=1):
    """get_child_elements.m:59-121 (method 1): per dimension the first child
    (lev+1, 2*pos) and, from level 1 on, the second (lev+1, 2*pos+1), when
    lev+1 < max_lev.  Returns (lev array, pos array) of the children in order."""
    if refinement_method != 1:
        raise NotImplementedError("only refinement method 1 is provided")
    lev_vec = list(int(v) for v in lev_vec)
    pos_vec = list(int(v) for v in pos_vec)
    levs, poss = [], []
    for d in range(len(lev_vec)):
        if lev_vec[d] + 1 < max_lev:
            l2, p2 = list(lev_vec), list(pos_vec)
            l2[d] += 1
            p2[d] *= 2
            levs.append(l2)
            poss.append(p2)
        if lev_vec[d] >= 1 and lev_vec[d] + 1 < max_lev:
            l2, p2 = list(lev_vec), list(pos_vec)
            l2[d] += 1
            p2[d] = 2 * p2[d] + 1
            levs.append(l2)
            poss.append(p2)
    nd = len(lev_vec)
    return np.asarray(levs, dtype=np.int64).reshape(-1, nd), np.asarray(poss, dtype=np.int64).reshape(-1, nd)


def lev_pos_to_id(lev, pos, max_lev):
    """md_lev_pos_to_idx.m vectorised over rows (1-based ids)."""
    lev = np.atleast_2d(lev)
    pos = np.atleast_2d(pos)
    idx = lev_cell_to_1d(lev, pos).astype(np.uint64)
    stride = np.uint64(1) << np.uint64(max_lev)
    acc = np.ones(lev.shape[0], dtype=np.uint64)
    mul = np.uint64(1)
    for d in range(lev.shape[1]):
        acc += (idx[:, d] - np.uint64(1)) * mul
        mul *= stride
    return acc


def refine_random(table, max_lev, keep_prob, rng):
    """One adaptive refinement sweep with a seeded mask: every method-1 child
    of every element that is not already present is kept with probability
    keep_prob; new elements are appended in parent order (the order adapt.m
    grows the element list).  Returns a new ElementTable (ids relative to max_lev)."""
    have = set(lev_pos_to_id(table.lev, table.cell, max_lev).tolist())
    new_lev, new_pos = [], []
    for e in range(table.num_elem):
        cl, cp = get_child_elements(table.lev[e], table.cell[e], max_lev)
        if cl.shape[0] == 0:
            continue
        ids = lev_pos_to_id(cl, cp, max_lev)
        keep = rng.random(cl.shape[0]) < keep_prob
        for k in range(cl.shape[0]):
            if keep[k] and int(ids[k]) not in have:
                have.add(int(ids[k]))
                new_lev.append(cl[k])
                new_pos.append(cp[k])
    if not new_lev:
        return table
    lev = np.vstack([table.lev, np.asarray(new_lev)])
    cell = np.vstack([table.cell, np.asarray(new_pos)])
    return ElementTable(lev, cell, lev_cell_to_1d(lev, cell), lev_pos_to_id(lev, cell, max_lev))
