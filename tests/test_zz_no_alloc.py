"""ocsl_vec_ABSsum_noalloc through the C ABI: the same result, bit for bit,
as the allocating ocsl_vec_ABSsum, on two calls in a row (the second with a
smaller input), with device inputs and output on the caller's stream, and no
change of the device's free memory across calls."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _abssum_ref(lib, a):
    """The allocating entry point on the same device input."""
    import torch

    torch.cuda.synchronize()  # its operands must be complete when it is called
    out = ctypes.c_double()
    st = lib.ocsl_vec_ABSsum(a.numel(), ctypes.c_void_p(a.data_ptr()), 1, ctypes.byref(out))
    assert st == 0, st
    return out.value


def test_noalloc_abssum_matches_allocating_entry_point():
    import torch

    from opencslsolver_amd._native import hip_lib
    from opencslsolver_amd.solver import _check

    lib = hip_lib()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(77)
    a1 = torch.from_numpy(rng.normal(size=3_000_003)).to(dev)
    a2 = torch.from_numpy(rng.normal(size=1_001)).to(dev)
    out = torch.full((2,), -1.0, dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def call(a, k):
        _check(lib.ocsl_vec_ABSsum_noalloc(a.numel(), ctypes.c_void_p(a.data_ptr()),
                                           ctypes.c_void_p(out[k:].data_ptr()), stream))

    call(a2, 0)  # first call on the device: looks up the scratch array's address
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    call(a1, 0)
    call(a2, 1)  # second call, smaller input, same scratch
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    got = out.cpu().numpy()
    r1, r2 = _abssum_ref(lib, a1), _abssum_ref(lib, a2)
    assert got[0] == r1 and got[1] == r2, (got, r1, r2)
    assert abs(r1 - np.abs(a1.cpu().numpy()).sum()) <= 1e-13 * r1
    assert free1 == free0
    # sizes beyond the maximum and null pointers are rejected with a message
    assert lib.ocsl_vec_ABSsum_noalloc(1 << 31, ctypes.c_void_p(a1.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()), stream) == -1
    assert b"2^31" in lib.ocsl_last_error()
    assert lib.ocsl_vec_ABSsum_noalloc(10, None, ctypes.c_void_p(out.data_ptr()), stream) == -1
