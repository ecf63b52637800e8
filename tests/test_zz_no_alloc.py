"""The allocation-free solve (taucs_amd_supernodal_solve_many_device_noalloc)
against the allocating one (taucs_amd_supernodal_solve_many_device), both
called through the C-ABI on the same factor and device buffers.

Two calls in a row, the second on a smaller input:
  1. mesh3d 48^3, 16 right-hand sides: 5.4 MB of workspace per right-hand
     side, so the 64 MiB scratch takes them in two sweeps (11 + 5);
  2. mesh3d 20^3, 4 right-hand sides: one sweep.
The library's device-allocation ledger must not move during either call.
Bar: FP64 solutions; a sweep of the same right-hand-side count runs the same
kernels with the same arithmetic, so the unsplit call must be bitwise equal;
the split call differs only in which small-front kernel variant a sweep uses,
so it must agree within 1e-12 relative per column (the allocating solve's own
multi- vs single-right-hand-side agreement is 1e-14).
"""
import numpy as np
import pytest

import taucs_amd as T

pytestmark = pytest.mark.gpu


def _case(side, nrhs, seed0):
    A = T.mesh3d(side)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    F = T.factor_llt_mf(PA)
    assert F is not None
    n = PA.n
    B = np.stack([T.random_uniform(n, seed0 + r) for r in range(nrhs)], 1)
    return PA, F, B


def _noalloc(F, dX, dB, n, nrhs):
    before = T.lib.taucs_amd_device_bytes_in_use()
    rc = T.lib.taucs_amd_supernodal_solve_many_device_noalloc(F.h, nrhs, dX.ptr, n, dB.ptr, n)
    after = T.lib.taucs_amd_device_bytes_in_use()
    assert rc == 0
    assert after == before, (before, after)
    return dX.get()


@pytest.mark.parametrize("first,second", [((48, 16), (20, 4))])
def test_noalloc_solve_matches_the_allocating_solve(first, second):
    T.lib.taucs_logfile(b"stderr")  # the library's messages, should a call fail
    assert T.lib.taucs_amd_noalloc_scratch_bytes() == 64 << 20
    for call, (side, nrhs) in enumerate((first, second)):
        print(f"case: noalloc call {call + 1}: mesh3d {side}^3, {nrhs} rhs", flush=True)
        PA, F, B = _case(side, nrhs, 300 + 40 * call)
        n = PA.n
        dB = T.DeviceArray(B)
        dX = T.DeviceArray(np.zeros_like(B))
        X0 = _noalloc(F, dX, dB, n, nrhs)
        dX2 = T.DeviceArray(np.zeros_like(B))
        F.solve_device(dX2.ptr, dB.ptr, n, nrhs)
        X1 = dX2.get()
        # and once more after the allocating solve ran (its workspace exists now)
        X2 = _noalloc(F, dX, dB, n, nrhs)
        assert np.array_equal(X0, X2)
        err = np.linalg.norm(X0 - X1, axis=0) / np.linalg.norm(X1, axis=0)
        print(f"  max relative difference to the allocating solve {err.max():.2e}", flush=True)
        assert err.max() <= 1e-12
        if call == 1:
            assert np.array_equal(X0, X1)
        M = PA.to_scipy_full()
        R = B - M @ X0
        bwd = abs(R).max(0) / (abs(M).sum(1).max() * abs(X0).max(0) + abs(B).max(0))
        assert bwd.max() <= 1e-14
        for d in (dB, dX, dX2):
            d.free()
        F.free()
