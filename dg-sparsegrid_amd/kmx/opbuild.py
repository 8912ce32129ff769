"""The 1D operators of a PDE built on the device (get_coeff_mats.m:29-53 through
kmx_sd_term_matrix): the pterm functions are sampled here at the library's
quadrature points and cell faces, the library assembles, solves, chains and
rotates.  Same terms as the CPU restatement in kmx.coeffs (the input provider
the parity tests pin against the reference's gold files)."""
import ctypes as C

import numpy as np

from . import _lib
from . import coeffs as K
from ._lib import check, lib

_TYPES = {"mass": _lib.PTERM_MASS, "div": _lib.PTERM_DIV, "grad": _lib.PTERM_GRAD}


def _filters(deg):
    H0, G0 = K.two_scale_filters(deg)
    return np.ascontiguousarray(H0, dtype=np.float64), np.ascontiguousarray(G0, dtype=np.float64)


def fmwt_matrix(deg, lev):
    """The forward multiwavelet transform (OperatorTwoScale_wavelet2.m) from the
    library, host, deg*2^lev square."""
    n = deg << lev
    H0, G0 = _filters(deg)
    out = np.empty(n * n)
    check(lib().kmx_fmwt_matrix(int(deg), int(lev), H0.ctypes.data_as(C.c_void_p), G0.ctypes.data_as(C.c_void_p),
                                out.ctypes.data_as(C.c_void_p)))
    return out.reshape(n, n).T.copy()  # column-major storage


def quad_points(deg, lev, xmin, xmax):
    """The library's quadrature points, (2^lev, nq)."""
    nq = C.c_int()
    check(lib().kmx_quad_points(int(deg), int(lev), float(xmin), float(xmax), C.byref(nq), None))
    x = np.empty((1 << lev) * nq.value)
    check(lib().kmx_quad_points(int(deg), int(lev), float(xmin), float(xmax), C.byref(nq),
                                x.ctypes.data_as(C.c_void_p)))
    return x.reshape(1 << lev, nq.value)


def _sample(fn, x, params, t):
    return np.ascontiguousarray(np.asarray(fn(x, params, t), dtype=np.float64) * np.ones_like(x), dtype=np.float64)


def sd_term_matrix(deg, t, dim, pterms, params, lev, stream=None):
    """sd_term_coeff_matrix.m for one dimension: the product of the pterms' M \\ L
    in the wavelet basis, built on the device.  Returns a (n, n) float64 CUDA
    tensor (a transposed view of the column-major result)."""
    import torch

    from .plan import _stream

    n = deg << lev
    xq = quad_points(deg, lev, dim.min, dim.max).ravel()
    h = (dim.max - dim.min) / (1 << lev)
    xf = dim.min + h * np.arange((1 << lev) + 1)
    keep = []
    descs = (_lib.PtermDesc * len(pterms))()
    for i, pt in enumerate(pterms):
        gq = _sample(pt.g, xq, params, t) * _sample(lambda x, p, tt, pt=pt: pt.dV(x, p, tt), xq, params, t)
        gf = _sample(pt.g, xf, params, t) * _sample(lambda x, p, tt, pt=pt: pt.dV(x, p, tt), xf, params, t)
        lq = _sample(pt.LHS_mass_g, xq, params, t) * _sample(lambda x, p, tt: dim.moment_dV(x), xq, params, t)
        keep += [gq, gf, lq]
        d = descs[i]
        d.type = _TYPES[pt.type]
        d.LF = float(pt.LF)
        d.ibcl = pt.IBCL.encode()
        d.ibcr = pt.IBCR.encode()
        d.g_quad = gq.ctypes.data_as(C.POINTER(C.c_double))
        d.g_face = gf.ctypes.data_as(C.POINTER(C.c_double))
        d.lhs_quad = lq.ctypes.data_as(C.POINTER(C.c_double))
    H0, G0 = _filters(deg)
    out = torch.empty(n * n, dtype=torch.float64, device="cuda")
    check(lib().kmx_sd_term_matrix(int(deg), int(lev), float(dim.min), float(dim.max), H0.ctypes.data_as(C.c_void_p),
                                   G0.ctypes.data_as(C.c_void_p), len(pterms), descs, C.c_void_p(out.data_ptr()),
                                   _stream(stream)))
    return out.view(n, n).t()


def get_coeff_mats(pde, deg, t=0.0, max_lev=None):
    """get_coeff_mats.m:29-53 on the device: mats[t][d] as (n, n) CUDA tensors
    (built at max_lev when given, as kmx.coeffs.get_coeff_mats)."""
    out = []
    for tm in pde.terms:
        row = []
        for d, dim in enumerate(pde.dims):
            lev = dim.lev if max_lev is None else max_lev
            row.append(sd_term_matrix(deg, t, dim, tm[d], pde.params, lev))
        out.append(row)
    return out
