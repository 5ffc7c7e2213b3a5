"""tcs_error_norms_noalloc (the allocation-free variant of tcs_error_norms)
through the C ABI: the same four results, bitwise, as tcs_error_norms on two
calls in a row, the second on a smaller grid (the static scratch keeps the
first call's partials, which the second must not use)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_error_norms_noalloc_equals_allocating_entry_point():
    import torch

    from tcstencil_amd import _lib

    assert torch.cuda.is_available()
    L = _lib.lib()
    rng = np.random.default_rng(71)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dev4 = torch.full((4,), -1.0, dtype=torch.float64, device="cuda")
    for rows, cols, r in ((900, 1300, 2), (130, 77, 1)):  # 2048 partial blocks, then 40
        print("case noalloc norms", rows, cols, r, flush=True)
        ref = torch.from_numpy(rng.standard_normal((rows, cols))).cuda()
        out = (ref + torch.from_numpy(rng.standard_normal((rows, cols)) * 1e-3).cuda()).float()
        host = np.zeros(4, dtype=np.float64)
        st = L.tcs_error_norms(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ref.data_ptr()), rows, cols, r,
                               host.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), stream)
        assert st == 0, L.tcs_last_error()
        dev4.fill_(-1.0)
        st = L.tcs_error_norms_noalloc(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ref.data_ptr()), rows,
                                       cols, r, ctypes.c_void_p(dev4.data_ptr()), stream)
        assert st == 0, L.tcs_last_error()
        torch.cuda.synchronize()
        got = dev4.cpu().numpy()
        assert np.array_equal(got, host), (got, host)
        d = (out.double() - ref)[r:rows - r, r:cols - r]
        np.testing.assert_allclose(got[1], float(d.norm() / ref[r:rows - r, r:cols - r].norm()), rtol=1e-12)
    # argument errors come back as status codes, before any launch
    assert L.tcs_error_norms_noalloc(None, None, 10, 10, 1, None, stream) == _lib.ERR_INVALID
    assert L.tcs_error_norms_noalloc(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ref.data_ptr()), 4, 4, 2,
                                     ctypes.c_void_p(dev4.data_ptr()), stream) == _lib.ERR_INVALID
