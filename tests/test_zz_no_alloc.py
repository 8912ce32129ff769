"""kmx_kronmult_noalloc (the library allocates nothing; intermediates in its fixed
scratch array) against the allocating kmx_kronmult, through the C ABI: two calls
in a row, the second smaller than the first (the scratch keeps the first call's
contents, which the second must not see), bit-for-bit equal results; an
intermediate beyond the scratch limit is refused with KMX_E_SHAPE."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(rng, shapes, nvec):
    import torch

    mats = [torch.from_numpy(rng.standard_normal(s)).cuda() for s in shapes]
    ncol = int(np.prod([s[1] for s in shapes]))
    X = torch.from_numpy(rng.standard_normal((ncol, nvec))).cuda()
    return mats, X


def test_kronmult_noalloc_matches_allocating_entry(kmx_gpu):
    import torch

    from kmx import _lib

    rng = np.random.default_rng(17)
    cases = [
        ([(4, 3), (5, 4), (3, 5), (4, 4)], 6),  # k = 4, rectangular factors, 6 vectors
        ([(3, 3), (2, 4), (3, 2)], 2),          # smaller: k = 3
        ([(4, 4), (4, 4)], 1),                  # k = 2, one vector
    ]
    for shapes, nvec in cases:
        mats, X = _case(rng, shapes, nvec)
        ref = kmx_gpu.kronmult(mats, X)
        got = kmx_gpu.kronmult(mats, X, noalloc=True)
        torch.cuda.synchronize()
        assert got.shape == ref.shape
        assert torch.equal(got, ref), (shapes, nvec, float((got - ref).abs().max()))
        # and the closed form, on the host
        K = mats[0].cpu().numpy()
        for m in mats[1:]:
            K = np.kron(K, m.cpu().numpy())
        want = K @ X.cpu().numpy()
        assert np.linalg.norm(got.cpu().numpy() - want) <= 1e-13 * np.linalg.norm(want)
    # an intermediate above KMX_KRON_SCRATCH_ELEMS / 2 values is refused
    big = [(128, 128), (128, 128), (128, 128)]
    mats, X = _case(rng, big, 1)
    assert 128 ** 3 > _lib.KRON_SCRATCH_ELEMS // 2
    with pytest.raises(kmx_gpu.KmxError) as e:
        kmx_gpu.kronmult(mats, X, noalloc=True)
    assert e.value.code == _lib.KMX_E_SHAPE
