"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/kmx.h declares, and its host-only entry points behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "kmx.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kmx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ["kmx_plan_create", "kmx_apply", "kmx_apply_transpose", "kmx_kronmult", "kmx_gemm_vbatched",
                 "kmx_kronmult_batch_plan", "kmx_batch_apply", "kmx_last_error", "kmx_plan_destroy"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import kmx._lib as L

    lib = L.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(declared_functions()) == set(L.PROTOTYPES), "binding prototypes out of sync with kmx.h"


def _replay(descs, work, mats, X):
    """Execute kmx_kronmult_batch_plan descriptors on the CPU (column-major buffers)."""
    k = len(mats)
    nvec = X.shape[1]
    Yn = int(np.prod([m.shape[0] for m in mats])) * nvec
    bufs = {0: np.asfortranarray(X).ravel(order="F").copy(), 1: np.zeros(Yn), 2: np.zeros(max(1, work)),
            3: np.zeros(1)}
    for i, m in enumerate(mats):
        bufs[4 + i] = np.asfortranarray(m).ravel(order="F").copy()
    stages = sorted({d["stage"] for d in descs})
    assert stages == list(range(1, k + 1))
    for d in descs:
        def view(b, off, rows, cols, ld):
            arr = bufs[b]
            return np.lib.stride_tricks.as_strided(arr[off:], shape=(rows, cols), strides=(8, 8 * ld))
        ra, ca = (d["k"], d["m"]) if d["transA"] == "T" else (d["m"], d["k"])
        rb, cb = (d["n"], d["k"]) if d["transB"] == "T" else (d["k"], d["n"])
        A = view(d["bufA"], d["offA"], ra, ca, d["lda"]).copy()
        B = view(d["bufB"], d["offB"], rb, cb, d["ldb"]).copy()
        A = A.T if d["transA"] == "T" else A
        B = B.T if d["transB"] == "T" else B
        Cv = view(d["bufC"], d["offC"], d["m"], d["n"], d["ldc"])
        Cv[...] = d["alpha"] * (A @ B) + (d["beta"] * Cv if d["beta"] != 0 else 0)
    return bufs[1].reshape((-1, nvec), order="F")


@pytest.mark.parametrize("shapes", [[(3, 3)], [(2, 3), (3, 4)], [(2, 3), (3, 4), (4, 5)], [(3, 3)] * 4,
                                    [(2, 3), (3, 4), (4, 5), (5, 6), (6, 7)],
                                    [(2, 3), (3, 4), (4, 5), (5, 6), (6, 7), (3, 2)], [(5, 2), (1, 6)]])
@pytest.mark.parametrize("nvec", [1, 2])
def test_kronmult_batch_plan_replays_to_kron(shapes, nvec):
    """Kronmult/test1_batch.m:33-221 pattern: the emitted lists reproduce kron(A1..Ak) X."""
    import kmx

    rng = np.random.default_rng(len(shapes) * 7 + nvec)
    mats = [rng.random(s) for s in shapes]
    X = rng.random((int(np.prod([s[1] for s in shapes])), nvec))
    descs, work = kmx.kronmult_batch_plan([s[0] for s in shapes], [s[1] for s in shapes], nvec)
    Y = _replay(descs, work, mats, X)
    K = mats[0]
    for m in mats[1:]:
        K = np.kron(K, m)
    assert np.allclose(Y, K @ X, rtol=1e-13, atol=1e-13)


def test_kronmult_batch_plan_square_counts():
    """Square deg x deg factors, one vector: stage 1 is one GEMM (kronmult1_batch on all
    deg^(k-1) columns), stage s >= 2 holds deg^(k-s) GEMMs (one per vector handed to that
    recursion level, kronmult{s}_batch.m); every stage costs 2 deg^(k+1) flops
    (the total flops_batch_list.m sums)."""
    import kmx

    deg, k = 3, 4
    descs, _ = kmx.kronmult_batch_plan([deg] * k, [deg] * k, 1)
    per_stage = [sum(1 for d in descs if d["stage"] == s) for s in range(1, k + 1)]
    assert per_stage == [1] + [deg ** (k - s) for s in range(2, k + 1)]
    flops = sum(2 * d["m"] * d["n"] * d["k"] for d in descs)
    assert flops == 2 * k * deg ** (k + 1)


def test_error_codes_host_side():
    import kmx._lib as L

    lib = L.lib()
    cnt = C.c_int()
    nr = (C.c_int * 2)(2, 0)
    rc = lib.kmx_kronmult_batch_plan(2, nr, nr, 1, None, 0, C.byref(cnt), None)
    assert rc == L.KMX_E_SHAPE
    assert b"positive" in lib.kmx_last_error()
    rc = lib.kmx_kronmult_batch_plan(7, nr, nr, 1, None, 0, C.byref(cnt), None)
    assert rc == L.KMX_E_ARG
    rc = lib.kmx_plan_create(None, None)
    assert rc == L.KMX_E_ARG
    with pytest.raises(L.KmxError):
        L.check(L.KMX_E_SHAPE)
