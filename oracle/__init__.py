"""CPU oracle for ASGarD-MATLAB's operator-application path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker (or the timed CPU
baseline).  The product (``dg-sparsegrid_amd/``) never imports it.

The numeric work lives in ``kmx_oracle.cpp`` (a C++17 FP64 restatement, built by
``oracle/Makefile`` into ``liboracle.so``); this module wraps it with ctypes and
returns numpy arrays in the reference's conventions (1-based element ids and 1D
indices, column-major 1D matrices, krond block order).

Parity status per function (what pins it; tests/test_oracle_gold.py):
  perm_eq_d / perm_leq_d / perm_eq / perm_leq / perm_max / index_leq_max
      gold/generated-inputs/permutations/*               -- pinned
  hash_table_sparse_nd / hash_table_nd / lev_cell_to_1d
      gold/generated-inputs/element_table/*              -- pinned
  connect_1d / connect_nd
      gold/generated-inputs/connectivity/*               -- pinned
  kron_multd / krond
      numpy.kron closed form on tiny rectangular inputs  -- pinned
  apply_A (both modes)
      assembled-A invariant (krond build path), 2D full-grid closed form,
      gold RK3 trajectories through the coefficient builder -- pinned
  gemm_vbatched
      per-entry closed form (numpy matmul)               -- pinned
  wavelet_to_realspace
      numpy.kron closed form summed over elements, and
      gold/generated-inputs/transformations/wavelet_to_realspace_continuity_{1,2,3}
      with the evaluation matrices of kmx.coeffs.matrix_plot_D (itself pinned by
      gold matrix_plot_D/*)                               -- pinned
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        L.orc_last_error.restype = C.c_char_p
        for name in ("orc_perm_eq_d", "orc_perm_leq_d", "orc_perm_max"):
            getattr(L, name).restype = C.c_long
        L.orc_perm_eq.restype = C.c_long
        L.orc_perm_leq.restype = C.c_long
        L.orc_index_leq_max.restype = C.c_long
        L.orc_hash_table_sparse_nd.restype = C.c_long
        L.orc_hash_table_nd.restype = C.c_long
        L.orc_connect_nd.restype = C.c_long
        L.orc_num_threads.restype = C.c_int
        _lib = L
        del i32p
    return _lib


def _chk(rc):
    if rc < 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())
    return rc


def _ip(a):
    return a.ctypes.data_as(C.c_void_p)


def _tuples(fn, ncol, *args):
    n = _chk(fn(*args, None))
    out = np.zeros((n, ncol), np.int32)
    _chk(fn(*args, _ip(out)))
    return out


# ------------------------------------------------------------------ PERMUTE/
def perm_eq_d(idim, levmax, sumeq, last_index_decreasing=False):
    lm = np.ascontiguousarray(np.broadcast_to(np.asarray(levmax, np.int32), (idim,)))
    return _tuples(lib().orc_perm_eq_d, idim, C.c_int(idim), _ip(lm), C.c_int(sumeq), C.c_int(int(last_index_decreasing)))


def perm_leq_d(idim, levmax, summax, increasing_sum_order=False):
    lm = np.ascontiguousarray(np.broadcast_to(np.asarray(levmax, np.int32), (idim,)))
    return _tuples(lib().orc_perm_leq_d, idim, C.c_int(idim), _ip(lm), C.c_int(summax), C.c_int(int(increasing_sum_order)))


def perm_eq(idim, n, last_index_decreasing=False):
    return _tuples(lib().orc_perm_eq, idim, C.c_int(idim), C.c_int(n), C.c_int(int(last_index_decreasing)))


def perm_leq(idim, n, order_by_n=False):
    return _tuples(lib().orc_perm_leq, idim, C.c_int(idim), C.c_int(n), C.c_int(int(order_by_n)))


def perm_max(idim, levmax, last_index_decreasing=False):
    lv = np.asarray(levmax, np.int32).ravel()
    if lv.size == 1:  # perm_max.m:23-26: scalar -> max in every dim
        lv = np.full(idim, lv.max(), np.int32)
    lv = np.ascontiguousarray(lv)
    return _tuples(lib().orc_perm_max, idim, C.c_int(idim), _ip(lv), C.c_int(int(last_index_decreasing)))


def index_leq_max(ndim, lists, levsum, levmax):
    vals = np.ascontiguousarray(np.concatenate([np.asarray(v, np.int32).ravel() for v in lists]))
    lens = np.asarray([len(np.ravel(v)) for v in lists], np.int32)
    return _tuples(lib().orc_index_leq_max, ndim, C.c_int(ndim), _ip(vals), _ip(lens), C.c_int(levsum), C.c_int(levmax))


# ------------------------------------------------------------ element table
def lev_cell_to_1d(lev, cell):
    return lib().orc_lev_cell_to_1d(C.c_int(lev), C.c_int(cell))


def hash_table_sparse_nd(lev_vec, max_lev, grid_type="SG"):
    """New hash.  Returns dict(lev, cell, idx1d: N x d int32, ids: uint64 1-based)."""
    lv = np.ascontiguousarray(np.asarray(lev_vec, np.int32).ravel())
    nd = lv.size
    sg = int(grid_type == "SG")
    L = lib()
    n = _chk(L.orc_hash_table_sparse_nd(nd, _ip(lv), max_lev, sg, None, None, None, None))
    lev = np.zeros((n, nd), np.int32)
    cell = np.zeros((n, nd), np.int32)
    idx = np.zeros((n, nd), np.int32)
    ids = np.zeros(n, np.uint64)
    _chk(L.orc_hash_table_sparse_nd(nd, _ip(lv), max_lev, sg, _ip(lev), _ip(cell), _ip(idx), _ip(ids)))
    return dict(lev=lev, cell=cell, idx1d=idx, ids=ids)


def hash_table_nd(nd, lev, grid_type="SG"):
    """Old hash (uniform level).  Returns dict(lev, cell, idx1d)."""
    L = lib()
    sg = int(grid_type == "SG")
    n = _chk(L.orc_hash_table_nd(nd, lev, sg, None, None, None))
    levo = np.zeros((n, nd), np.int32)
    cell = np.zeros((n, nd), np.int32)
    idx = np.zeros((n, nd), np.int32)
    _chk(L.orc_hash_table_nd(nd, lev, sg, _ip(levo), _ip(cell), _ip(idx)))
    return dict(lev=levo, cell=cell, idx1d=idx)


def connect_1d(lev):
    n = 1 << lev
    out = np.zeros((n, n), np.uint8)
    _chk(lib().orc_connect_1d(lev, _ip(out)))
    return out


def connect_nd(nd, lev, grid_type_table, levsum, levmax, sg_filter=True):
    """connect_nD over the old hash of (nd, lev, grid).  Returns (row_ptr int64 0-based, cols int32 1-based)."""
    L = lib()
    sgt = int(grid_type_table == "SG")
    N = _chk(L.orc_hash_table_nd(nd, lev, sgt, None, None, None))
    rp = np.zeros(N + 1, np.int64)
    P = _chk(L.orc_connect_nd(nd, lev, sgt, levsum, levmax, int(sg_filter), _ip(rp), None))
    cols = np.zeros(P, np.int32)
    _chk(L.orc_connect_nd(nd, lev, sgt, levsum, levmax, int(sg_filter), _ip(rp), _ip(cols)))
    return rp, cols


# ------------------------------------------------------------------- kron
def _matlist(mats):
    mats = [np.asfortranarray(np.asarray(m, np.float64)) for m in mats]
    ptrs = (C.c_void_p * len(mats))(*[m.ctypes.data for m in mats])
    nrow = np.asarray([m.shape[0] for m in mats], np.int32)
    ncol = np.asarray([m.shape[1] for m in mats], np.int32)
    return mats, ptrs, nrow, ncol


def kron_multd(mats, X):
    """Y = kron(A1..Ak) X through the reference's kronmult recursion.  X: (prod ncol) x nvec (Fortran order)."""
    keep, ptrs, nrow, ncol = _matlist(mats)
    X = np.asarray(X, np.float64)
    isx = int(np.prod(ncol))
    Xf = np.ascontiguousarray(X.reshape(-1, order="F"))
    nvec = Xf.size // isx
    Y = np.zeros(int(np.prod(nrow)) * nvec, np.float64)
    _chk(lib().orc_kron_multd(len(keep), _ip(nrow), _ip(ncol), ptrs, _ip(Xf), C.c_long(nvec), _ip(Y)))
    return Y.reshape((-1, nvec), order="F")


def krond(mats):
    keep, ptrs, nrow, ncol = _matlist(mats)
    Y = np.zeros(int(np.prod(nrow)) * int(np.prod(ncol)), np.float64)
    _chk(lib().orc_krond(len(keep), _ip(nrow), _ip(ncol), ptrs, _ip(Y)))
    return Y.reshape((int(np.prod(nrow)), int(np.prod(ncol))), order="F")


def wavelet_to_realspace(meval, idx1d, deg, fw):
    """wavelet_to_realspace.m:44-79: sum over elements of kron(Meval_d(:, block of
    I_d(e))) * f_w(block e).  meval: list of npts_d x ncols_d arrays; idx1d: N x d
    1-based; returns prod(npts) values in Kronecker order (dim 1 slowest)."""
    keep, ptrs, nrow, ncol = _matlist(meval)
    idx = np.ascontiguousarray(np.asarray(idx1d, np.int32).reshape(-1, len(keep)))
    fw = np.ascontiguousarray(np.asarray(fw, np.float64))
    fr = np.zeros(int(np.prod(nrow)), np.float64)
    _chk(lib().orc_wavelet_to_realspace(len(keep), int(deg), idx.shape[0], _ip(idx), _ip(nrow), _ip(ncol), ptrs,
                                        _ip(fw), _ip(fr)))
    return fr


class ApplyInputs:
    """Operator inputs in the reference's conventions.

    mats[t][d]: dense n_d x n_d FP64 (pde.terms{t}.terms_1D{d}.mat);
    idx1d: N x d 1-based (A_data.element_local_index_D);
    row_ptr/cols: connectivity as CSR (cols 1-based) or None for all pairs."""

    def __init__(self, mats, idx1d, deg, row_ptr=None, cols=None):
        self.T = len(mats)
        self.ndim = len(mats[0])
        self.deg = int(deg)
        self.idx1d = np.ascontiguousarray(np.asarray(idx1d, np.int32))
        self.N = self.idx1d.shape[0]
        self._mats = [np.asfortranarray(np.asarray(m, np.float64)) for tm in mats for m in tm]
        self._ptrs = (C.c_void_p * len(self._mats))(*[m.ctypes.data for m in self._mats])
        self.n = np.asarray([m.shape[0] for m in self._mats[: self.ndim]], np.int32)
        self.row_ptr = None if row_ptr is None else np.ascontiguousarray(np.asarray(row_ptr, np.int64))
        self.cols = None if cols is None else np.ascontiguousarray(np.asarray(cols, np.int32))
        self.B = self.deg ** self.ndim

    def _conn(self):
        if self.row_ptr is None:
            return None, None
        return _ip(self.row_ptr), _ip(self.cols)


def apply_A(inp, x, mode=0, nthreads=1):
    """y = A x exactly as apply_A.m's explicit branch (elem -> j -> t)."""
    x = np.ascontiguousarray(np.asarray(x, np.float64))
    y = np.zeros(inp.N * inp.B, np.float64)
    rp, cc = inp._conn()
    _chk(lib().orc_apply_A(inp.ndim, inp.deg, inp.N, inp.T, _ip(inp.n), inp._ptrs, _ip(inp.idx1d), rp, cc,
                           _ip(x), _ip(y), mode, nthreads))
    return y


def build_A(inp):
    """Dense assembled A (apply_A.m build_A branch)."""
    dof = inp.N * inp.B
    A = np.zeros(dof * dof, np.float64)
    rp, cc = inp._conn()
    _chk(lib().orc_build_A(inp.ndim, inp.deg, inp.N, inp.T, _ip(inp.n), inp._ptrs, _ip(inp.idx1d), rp, cc, _ip(A)))
    return A.reshape((dof, dof), order="F")


def gemm_vbatched(transA, transB, ms, ns, ks, alpha, As, Bs, beta, Cs):
    """Reference gemm_vbatched (k from klist).  Cs are updated in place (Fortran arrays)."""
    cnt = len(ms)
    As = [np.asfortranarray(a, np.float64) for a in As]
    Bs = [np.asfortranarray(b, np.float64) for b in Bs]
    for c in Cs:
        assert c.flags.f_contiguous and c.dtype == np.float64
    lda = np.asarray([a.shape[0] for a in As], np.int32)
    ldb = np.asarray([b.shape[0] for b in Bs], np.int32)
    ldc = np.asarray([c.shape[0] for c in Cs], np.int32)
    pa = (C.c_void_p * cnt)(*[a.ctypes.data for a in As])
    pb = (C.c_void_p * cnt)(*[b.ctypes.data for b in Bs])
    pc = (C.c_void_p * cnt)(*[c.ctypes.data for c in Cs])
    m = np.asarray(ms, np.int32)
    n = np.asarray(ns, np.int32)
    k = np.asarray(ks, np.int32)
    _chk(lib().orc_gemm_vbatched(C.c_char(transA.encode()), C.c_char(transB.encode()), _ip(m), _ip(n), _ip(k),
                                 C.c_double(alpha), pa, _ip(lda), pb, _ip(ldb), C.c_double(beta), pc, _ip(ldc),
                                 C.c_int(cnt)))
    return Cs


def num_threads():
    return lib().orc_num_threads()
