"""The allocation-free tall-skinny product (qmbed_hip_gemm_atb_noalloc)
against its allocating counterpart (qmbed_hip_gemm_atb), through the C ABI:
two calls in a row, the second on a smaller input, bit for bit equal; a shape
beyond the static scratch is rejected before any launch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _product(qa, fn, A, B):
    p, n = A.shape
    q = B.shape[0]
    bufs = [qa.DeviceBuffer(max(x.nbytes, 8)) for x in (A, B)] + [qa.DeviceBuffer(max(p * q * 8, 8))]
    bufs[0].upload(A)
    bufs[1].upload(B)
    f = getattr(qa.library(), fn)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    rc = f(bufs[0].ptr, n, p, bufs[1].ptr, q, bufs[2].ptr)
    C = bufs[2].download(np.empty(p * q)).reshape(q, p).T if rc == 0 else None
    for b in bufs:
        b.free()
    return rc, C


def test_gemm_atb_noalloc_matches_allocating_entry(qa):
    rng = np.random.default_rng(11)
    for (n, p, q) in ((300001, 64, 64), (4097, 33, 5)):  # a large call, then a smaller one
        print(f"[no-alloc] n={n} p={p} q={q}", flush=True)
        A = rng.standard_normal((p, n))
        B = rng.standard_normal((q, n))
        rc1, C1 = _product(qa, "qmbed_hip_gemm_atb", A, B)
        rc2, C2 = _product(qa, "qmbed_hip_gemm_atb_noalloc", A, B)
        assert rc1 == 0 and rc2 == 0
        assert np.array_equal(C1, C2)
        np.testing.assert_allclose(C2, A @ B.T, rtol=1e-10, atol=1e-9 * np.sqrt(n))
    # a shape whose partial products exceed the 4 MiB scratch: rejected
    A = rng.standard_normal((1024, 70))
    rc, _ = _product(qa, "qmbed_hip_gemm_atb_noalloc", A, A)
    assert rc == -2
    # the allocating entry point takes it
    rc, C = _product(qa, "qmbed_hip_gemm_atb", A, A)
    assert rc == 0
    np.testing.assert_allclose(C, A @ A.T, rtol=1e-10, atol=1e-9)
