"""The no-allocation dense solve (spx_solve_batch_f64_device_noalloc): same
results as spx_solve_batch_f64_device, bit for bit, on two calls in a row
(the second with a smaller batch, so the scratch array still holds the first
call's systems), on the register path (n <= 64) and on the generic path
whose LU workspace is the library's static scratch array; sizes past that
array are rejected."""
import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import engine as E

pytestmark = pytest.mark.gpu


def _systems(n, B, seed):
    rng = np.random.default_rng(seed)
    M = rng.uniform(-1.0, 1.0, (B, n, n)) + n * np.eye(n)
    M[::7] *= rng.uniform(0.5, 2.0)           # vary the scale between systems
    b = rng.uniform(-1.0, 1.0, (B, n))
    return M, b


def _solve(torch, lib, handle, fn, M, b, n):
    B = len(M)
    dev = torch.device("cuda:0")
    dM = torch.as_tensor(M.transpose(0, 2, 1).copy(), device=dev)
    db = torch.as_tensor(b, device=dev)
    dx = torch.full((B, n), np.nan, dtype=torch.float64, device=dev)
    ds = torch.full((B,), 7, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    rc = getattr(lib, fn)(handle, A.ptr(dM), A.ptr(db), n, B, A.ptr(dx), A.ptr(ds))
    assert rc == A.SPX_OK, E.Device.get(0).lib.spx_last_error()
    assert lib.spx_synchronize(handle) == A.SPX_OK
    return dx.cpu().numpy(), ds.cpu().numpy()


@pytest.mark.parametrize("n,sizes", [(40, (3000, 700)), (100, (3000, 700)), (700, (40, 9))])
def test_noalloc_solve_equals_allocating_solve(n, sizes):
    import torch
    dev = E.Device.get(0)
    lib = dev.lib
    for call, B in enumerate(sizes):
        M, b = _systems(n, B, seed=10 * n + call)
        if call == 1:
            M[3] = 0.0  # a singular system in the second call
        want_x, want_s = _solve(torch, lib, dev.handle, "spx_solve_batch_f64_device", M, b, n)
        got_x, got_s = _solve(torch, lib, dev.handle, "spx_solve_batch_f64_device_noalloc", M,
                              b, n)
        np.testing.assert_array_equal(got_s, want_s)
        np.testing.assert_array_equal(got_x, want_x)  # the same kernels: bit for bit
        assert (got_s[3] == 1) == (call == 1)
        ok = got_s == 0
        r = np.einsum("bij,bj->bi", M[ok], got_x[ok]) - b[ok]
        assert np.abs(r).max() <= 1e-10 * n


def test_noalloc_solve_rejects_systems_past_the_scratch_array():
    dev = E.Device.get(0)
    lib = dev.lib
    one = np.zeros(1)
    p = one.ctypes.data
    # n = 2896 needs more than the 64 MiB array for one system: rejected
    # before any device access
    rc = lib.spx_solve_batch_f64_device_noalloc(dev.handle, p, p, 2896, 1, p, None)
    assert rc == A.SPX_ERR_INVALID_DIMENSION
    assert b"2895" in lib.spx_last_error()
    assert lib.spx_solve_batch_f64_device_noalloc(dev.handle, p, p, 8, 0, p, None) == A.SPX_OK
