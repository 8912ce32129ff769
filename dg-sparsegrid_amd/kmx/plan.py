"""Python host side of the C ABI: resident operator plans, batches, and the
standalone Kronecker entry points.

Device memory and streams come from PyTorch (plumbing only): vectors are
``torch.float64`` CUDA tensors and launches go on torch's current stream.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import KmxError, check, lib  # noqa: F401

PREC = {
    "fp64": _lib.PREC_FP64,
    "fp32": _lib.PREC_FP32,
    "bf16": _lib.PREC_BF16,
}


def _prec(p):
    if isinstance(p, str):
        return PREC[p]
    return int(p)


_TORCH = None  # torch, once a device was seen (the check is not repeated per call)


def _torch():
    global _TORCH
    if _TORCH is None:
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("kmx: no HIP device visible to torch")
        _TORCH = torch
    return _TORCH


def _stream(stream):
    if stream is None:
        return C.c_void_p(_torch().cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def _ptr(t):
    """Device pointer of a contiguous float64 CUDA tensor (or pass an int)."""
    if isinstance(t, int):
        return C.c_void_p(t)
    if not (t.is_cuda and t.dtype == _torch().float64 and t.is_contiguous()):
        raise ValueError("kmx: vectors must be contiguous float64 CUDA tensors")
    return C.c_void_p(t.data_ptr())


class Plan:
    """A resident operator y = sum_{(i,j)} sum_t kron_d A_{t,d}[I_d(i), J_d(j)] x_j.

    mats      list (terms) of lists (dims) of square n_d x n_d arrays
              (pde.terms{t}.terms_1D{d}.mat)
    idx1d     N x d 1D indices per element (A_data.element_local_index_D)
    deg       block edge
    row_ptr   CSR offsets (N+1, 0-based) of the pair list, or None = all pairs
    cols      connected element ids
    index_base 1 if idx1d / cols are 1-based (MATLAB), 0 otherwise
    fp32      keep FP32 block copies (KMX_PLAN_FP32)
    adjoint   keep the transposed pair list (KMX_PLAN_ADJOINT)
    level_blocks  also detect the level-block structure (KMX_PLAN_LEVEL_BLOCKS)
              and build the dimension sweep on it (set_sweep / autotune; BF16
              applies need its separable form); off, every apply runs pair by pair
    identity  None, or T x d booleans declaring factors the identity (MASS
              pterms with g = 1): the separable sweep takes them as exactly I
              (kmx_plan_desc.identity)
    """

    def __init__(self, mats, idx1d, deg, row_ptr=None, cols=None, index_base=1, fp32=True, adjoint=False,
                 level_blocks=False, identity=None):
        _torch().cuda.current_device()  # make sure the HIP context exists
        self.T = len(mats)
        self.ndim = len(mats[0])
        self.deg = int(deg)
        idx = np.ascontiguousarray(np.asarray(idx1d, dtype=np.int32).reshape(-1, self.ndim))
        self.N = idx.shape[0]
        self.B = self.deg ** self.ndim
        keep = [np.asfortranarray(np.asarray(m, dtype=np.float64)) for tm in mats for m in tm]
        if len(keep) != self.T * self.ndim:
            raise ValueError("every term needs one matrix per dimension")
        rows = np.asarray([keep[d].shape[0] for d in range(self.ndim)], dtype=np.int32)
        for t in range(self.T):
            for d in range(self.ndim):
                m = keep[t * self.ndim + d]
                if m.shape != (rows[d], rows[d]):
                    raise ValueError(f"term {t} dim {d}: matrix {m.shape} != ({rows[d]}, {rows[d]})")
        ptrs = (C.c_void_p * len(keep))(*[m.ctypes.data for m in keep])
        desc = _lib.PlanDesc()
        desc.ndim = self.ndim
        desc.deg = self.deg
        desc.num_elem = self.N
        desc.num_terms = self.T
        desc.mat_rows = rows.ctypes.data_as(C.POINTER(C.c_int))
        desc.mats = C.cast(ptrs, C.POINTER(C.c_void_p))
        desc.elem_idx1d = idx.ctypes.data_as(C.POINTER(C.c_int))
        if row_ptr is not None:
            rp = np.ascontiguousarray(np.asarray(row_ptr, dtype=np.int64))
            cc = np.ascontiguousarray(np.asarray(cols, dtype=np.int32))
            desc.row_ptr = rp.ctypes.data_as(C.POINTER(C.c_int64))
            desc.cols = cc.ctypes.data_as(C.POINTER(C.c_int))
        desc.index_base = int(index_base)
        desc.flags = ((_lib.PLAN_FP32 if fp32 else 0) | (_lib.PLAN_ADJOINT if adjoint else 0)
                      | (_lib.PLAN_LEVEL_BLOCKS if level_blocks else 0))
        if identity is not None:
            ident = np.ascontiguousarray(np.asarray(identity, dtype=np.uint8).reshape(self.T * self.ndim))
            desc.identity = ident.ctypes.data_as(C.POINTER(C.c_ubyte))
        h = C.c_void_p()
        check(lib().kmx_plan_create(C.byref(desc), C.byref(h)))
        self._h = h
        self.mat_rows = rows
        P = C.c_int64()
        check(lib().kmx_plan_info(h, C.byref(P), None, None))
        self.P = P.value
        self.dof = self.N * self.B

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("kmx: plan destroyed")
        return self._h

    def apply(self, x, y=None, prec="fp64", stream=None):
        """y = A x (overwrites y; allocates it when None)."""
        torch = _torch()
        if y is None:
            y = torch.empty(self.dof, dtype=torch.float64, device=x.device)
        self._check_len(x, y)
        check(lib().kmx_apply(self.handle, _ptr(x), _ptr(y), _prec(prec), _stream(stream)))
        return y

    def apply_transpose(self, x, y=None, prec="fp64", stream=None):
        """y = A^T x (needs adjoint=True)."""
        torch = _torch()
        if y is None:
            y = torch.empty(self.dof, dtype=torch.float64, device=x.device)
        self._check_len(x, y)
        check(lib().kmx_apply_transpose(self.handle, _ptr(x), _ptr(y), _prec(prec), _stream(stream)))
        return y

    def assemble(self, stream=None):
        """The assembled operator's blocks (kmx_plan_assemble; apply_A.m's build_A
        branch): a (P, B, B) FP64 device tensor, blocks[q, c, r] = entry (r, c) of
        pair q's block (column-major per block, A_s3's order)."""
        torch = _torch()
        out = torch.empty((self.P, self.B, self.B), dtype=torch.float64, device="cuda")
        check(lib().kmx_plan_assemble(self.handle, _ptr(out) if self.P else C.c_void_p(0), _stream(stream)))
        return out

    def apply_host(self, x, prec="fp64"):
        """numpy in, numpy out (copies through the library's staging buffers)."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        if x.size != self.dof:
            raise ValueError(f"x has {x.size} entries, plan expects {self.dof}")
        y = np.empty_like(x)
        check(lib().kmx_apply_host(self.handle, x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p),
                                   _prec(prec)))
        return y

    def update_term(self, t, d, mat, stream=None):
        m = np.asfortranarray(np.asarray(mat, dtype=np.float64))
        if m.shape != (self.mat_rows[d], self.mat_rows[d]):
            raise ValueError("update_term: wrong matrix size")
        check(lib().kmx_plan_update_term(self.handle, int(t), int(d), m.ctypes.data_as(C.c_void_p), _stream(stream)))

    def autotune(self, stream=None):
        """Time the kernel variants on this plan and keep the fastest (kmx_plan_autotune)."""
        check(lib().kmx_plan_autotune(self.handle, _stream(stream)))
        return {p: dict(self.variant(p), sweep=self.sweep(p)) for p in ("fp64", "fp32")}

    def variant(self, prec="fp64"):
        nt, lds = C.c_int(), C.c_int()
        check(lib().kmx_plan_get_variant(self.handle, _prec(prec), C.byref(nt), C.byref(lds)))
        return {"threads": nt.value, "lds_staged": bool(lds.value), "lds_mode": lds.value}

    def set_sweep(self, prec, on=True):
        """Use (or stop using) the dimension-sweep path for prec (kmx_plan_set_sweep)."""
        check(lib().kmx_plan_set_sweep(self.handle, _prec(prec), int(bool(on))))

    def sweep(self, prec="fp64"):
        av, ac, fl = C.c_int(), C.c_int(), C.c_double()
        check(lib().kmx_plan_get_sweep(self.handle, _prec(prec), C.byref(av), C.byref(ac), C.byref(fl)))
        return {"available": bool(av.value), "active": bool(ac.value), "flops": fl.value}

    def sweep_info(self):
        """The dimension sweep's form and storage (kmx_plan_sweep_info)."""
        info = _lib.SweepInfo()
        check(lib().kmx_plan_sweep_info(self.handle, C.byref(info)))
        out = {f: getattr(info, f) for f, _ in _lib.SweepInfo._fields_}
        for f in ("available", "separable", "fp64"):
            out[f] = bool(out[f])
        return out

    def set_term_scale(self, t, alpha):
        check(lib().kmx_plan_set_term_scale(self.handle, int(t), float(alpha)))

    def _check_len(self, x, y):
        if x.numel() != self.dof or y.numel() != self.dof:
            raise ValueError(f"vectors must have {self.dof} entries")

    def destroy(self):
        if getattr(self, "_h", None) is not None:
            lib().kmx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Several independent plans applied in one persistent launch."""

    def __init__(self, plans):
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        h = C.c_void_p()
        check(lib().kmx_batch_create(arr, len(self.plans), C.byref(h)))
        self._h = h

    def set_sweep(self, prec, on=True):
        """Run the plans that have a dimension-sweep path as merged sweeps, one
        launch per phase for all of them (kmx_batch_set_sweep); not under capture."""
        check(lib().kmx_batch_set_sweep(self._h, _prec(prec), int(bool(on))))

    def sweep(self, prec="fp32"):
        m, a, ph = C.c_int(), C.c_int(), C.c_int()
        check(lib().kmx_batch_get_sweep(self._h, _prec(prec), C.byref(m), C.byref(a), C.byref(ph)))
        return {"members": m.value, "active": bool(a.value), "phases": ph.value}

    def autotune(self, prec="fp32", stream=None):
        """Keep the fastest load batches per wave of the merged sweep for prec
        (kmx_batch_autotune); returns the choice.  Not under capture."""
        wb = C.c_int()
        check(lib().kmx_batch_autotune(self._h, _prec(prec), C.byref(wb), _stream(stream)))
        return wb.value

    def apply(self, xs, ys, prec="fp32", stream=None):
        n = len(self.plans)
        px = (C.c_void_p * n)(*[_ptr(x).value for x in xs])
        py = (C.c_void_p * n)(*[_ptr(y).value for y in ys])
        check(lib().kmx_batch_apply(self._h, px, py, _prec(prec), _stream(stream)))
        return ys

    def rk3(self, fs, dts, nsteps=1, prec="fp32", stream=None):
        """nsteps fused RK3 steps of every plan (kmx_batch_rk3): fs advanced in place."""
        n = len(self.plans)
        pf = (C.c_void_p * n)(*[_ptr(f).value for f in fs])
        pd = (C.c_double * n)(*[float(d) for d in dts])
        check(lib().kmx_batch_rk3(self._h, pf, pd, int(nsteps), _prec(prec), _stream(stream)))
        return fs

    def destroy(self):
        if getattr(self, "_h", None) is not None:
            lib().kmx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class RealSpace:
    """Wavelet space -> real space (wavelet_to_realspace.m): f_r = sum_e
    kron_d Meval_d(:, block J_d(e)) f_w(block e), on the device.

    meval     list (dims) of npts_d x ncols_d arrays (Meval_D{d}, host)
    idx1d     N x d 1D indices per element; deg block edge"""

    def __init__(self, meval, idx1d, deg, index_base=1):
        _torch().cuda.current_device()
        keep = [np.asfortranarray(np.asarray(m, dtype=np.float64)) for m in meval]
        self.ndim = len(keep)
        self.deg = int(deg)
        idx = np.ascontiguousarray(np.asarray(idx1d, dtype=np.int32).reshape(-1, self.ndim))
        self.N = idx.shape[0]
        npts = np.asarray([m.shape[0] for m in keep], dtype=np.int32)
        ncols = np.asarray([m.shape[1] for m in keep], dtype=np.int32)
        ptrs = (C.c_void_p * self.ndim)(*[m.ctypes.data for m in keep])
        h = C.c_void_p()
        check(lib().kmx_realspace_create(self.ndim, self.deg, self.N, idx.ctypes.data_as(C.POINTER(C.c_int)),
                                         int(index_base), npts.ctypes.data_as(C.POINTER(C.c_int)),
                                         ncols.ctypes.data_as(C.POINTER(C.c_int)),
                                         C.cast(ptrs, C.POINTER(C.c_void_p)), C.byref(h)))
        self._h = h
        self.npts = [int(v) for v in npts]
        self.dof = self.N * self.deg ** self.ndim

    def apply(self, f_w, f_r=None, stream=None):
        torch = _torch()
        if f_w.numel() != self.dof:
            raise ValueError(f"f_w must have {self.dof} entries")
        if f_r is None:
            f_r = torch.empty(int(np.prod(self.npts)), dtype=torch.float64, device=f_w.device)
        check(lib().kmx_realspace_apply(self._h, _ptr(f_w), _ptr(f_r), _stream(stream)))
        return f_r

    def destroy(self):
        if getattr(self, "_h", None) is not None:
            lib().kmx_realspace_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def arnoldi_step(n, k, V, u, s0, s1, h, work, stream=None):
    """One Arnoldi step on device tensors (kmx_arnoldi_step): w = s0 V[k] + s1 u
    orthogonalised against V[0..k] (classical Gram-Schmidt twice), h[0..k+1] its
    coefficients and norm, V[k+1] = w / |w|.  V: (>= k+2, n) float64, u: n,
    h: >= k+2, work: kmx._lib.arnoldi_work(n) float64, all CUDA tensors."""
    if V.dim() != 2 or V.shape[1] != n or V.shape[0] < k + 2 or not V.is_contiguous():
        raise ValueError("arnoldi_step: V must be a contiguous (>= k+2, n) tensor")
    if u.numel() != n or h.numel() < k + 2 or work.numel() < _lib.arnoldi_work(n):
        raise ValueError("arnoldi_step: u, h or work too small")
    check(lib().kmx_arnoldi_step(int(n), int(k), _ptr(V), _ptr(u), float(s0), float(s1), _ptr(h), _ptr(work),
                                 _stream(stream)))


def kronmult(mats, X, stream=None, noalloc=False, Y=None):
    """Y = kron(A1..Ak) X on the device (kron_multd.m semantics, k = 1..16).

    mats: 2-D float64 CUDA tensors (nrow_i, ncol_i); X: (prod ncol, nvec)
    float64 CUDA tensor (or 1-D for nvec = 1).  Returns Y (prod nrow, nvec).
    Column-major copies are made for the C ABI, which uses MATLAB layout.
    noalloc: call kmx_kronmult_noalloc (the library allocates nothing; its
    intermediates use the library's fixed scratch array)."""
    torch = _torch()
    k = len(mats)
    cm = [m.t().contiguous() for m in mats]  # storage = column-major of m
    nr = [int(m.shape[0]) for m in mats]
    nc = [int(m.shape[1]) for m in mats]
    ncx = int(np.prod(nc))
    X2 = X.reshape(ncx, -1)
    nvec = X2.shape[1]
    xs = X2.t().contiguous()  # column-major storage of X
    y = torch.empty(int(np.prod(nr)) * nvec, dtype=torch.float64, device=X.device) if Y is None else Y
    fn = lib().kmx_kronmult_noalloc if noalloc else lib().kmx_kronmult
    check(fn(k, (C.c_int * k)(*nr), (C.c_int * k)(*nc), (C.c_void_p * k)(*[m.data_ptr() for m in cm]),
             C.c_void_p(xs.data_ptr()), int(nvec), C.c_void_p(y.data_ptr()), _stream(stream)))
    return y.view(nvec, -1).t()


def gemm_vbatched(transA, transB, ms, ns, ks, alpha, As, ldas, Bs, ldbs, beta, Cs, ldcs, stream=None):
    """Batched GEMM over device pointers (ints or tensors)."""
    cnt = len(ms)

    def arr(v):
        return (C.c_int * cnt)(*[int(a) for a in v])

    def parr(v):
        return (C.c_void_p * cnt)(*[(a if isinstance(a, int) else a.data_ptr()) for a in v])

    check(lib().kmx_gemm_vbatched(transA.encode(), transB.encode(), arr(ms), arr(ns), arr(ks), float(alpha),
                                  parr(As), arr(ldas), parr(Bs), arr(ldbs), float(beta), parr(Cs), arr(ldcs), cnt,
                                  _stream(stream)))


def kronmult_batch_plan(nrow, ncol, nvec):
    """Descriptor lists of Kronmult/kronmult{k}_batch.m (host only).
    Returns (list of dicts in stage-major order, workspace elements)."""
    k = len(nrow)
    nr = (C.c_int * k)(*[int(v) for v in nrow])
    nc = (C.c_int * k)(*[int(v) for v in ncol])
    cnt = C.c_int()
    work = C.c_int64()
    check(lib().kmx_kronmult_batch_plan(k, nr, nc, int(nvec), None, 0, C.byref(cnt), C.byref(work)))
    out = (_lib.GemmDesc * max(1, cnt.value))()
    check(lib().kmx_kronmult_batch_plan(k, nr, nc, int(nvec), out, cnt.value, C.byref(cnt), C.byref(work)))
    descs = []
    for g in out[: cnt.value]:
        descs.append({f: getattr(g, f) for f, _ in _lib.GemmDesc._fields_})
        descs[-1]["transA"] = descs[-1]["transA"].decode()
        descs[-1]["transB"] = descs[-1]["transB"].decode()
    return descs, work.value


def version():
    return lib().kmx_version().decode()
