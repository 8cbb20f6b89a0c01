"""The allocation-free single-matrix entry point
(lsvd_small_solve_rectangular_f64_noalloc) against the allocating one
(lsvd_small_solve_rectangular_f64), both through the C ABI on device
buffers.

Two calls in a row on one handle, the second with a smaller input, so the
second runs on the scratch array the first left behind (queue counter and
staging slot reused).  Bar: the FP64 two-tier sigma parity (relative 1e-12
above 1e-3 sigma_max, absolute 64 n eps sigma_max), singular-vector parity
per cluster, the same info counters, and the reference's residual checks;
the two entry points run the same kernel variant, so they are expected to
agree to the last bit, which is checked too.  The error paths (host
pointer mode, a size above the scratch's limit) are checked without a
launch.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from lie_cliffalg_analog_svd_amd import _lib
from lie_cliffalg_analog_svd_amd.handle import Handle
from tests._util import assert_factors, assert_sigma_close, uniform_matrix, vector_parity

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _solve(h, fn, a):
    import torch
    m, n = a.shape
    k = min(m, n)
    dev = torch.device("cuda:0")
    da = torch.from_numpy(a).to(dev)
    du = torch.full((m, k), np.nan, dtype=torch.float64, device=dev)
    ds = torch.full((k,), np.nan, dtype=torch.float64, device=dev)
    dvt = torch.full((k, n), np.nan, dtype=torch.float64, device=dev)
    dinfo = torch.zeros(C.sizeof(_lib.LsvdInfo) // 4, dtype=torch.int32, device=dev)
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    torch.cuda.synchronize()
    h.check(fn(h.raw, p(da), m, n, p(du), p(ds), p(dvt), p(dinfo)))
    h.synchronize()
    info = _lib.LsvdInfo.from_buffer_copy(dinfo.cpu().numpy().tobytes()).as_dict()
    return du.cpu().numpy(), ds.cpu().numpy(), dvt.cpu().numpy(), info


@pytest.mark.parametrize("shapes", [((128, 100), (33, 20)), ((96, 96), (70, 64)), ((40, 120), (5, 3))])
def test_noalloc_matches_allocating_entry_point(shapes):
    L = _lib.lib()
    h = Handle()
    h.set_pointer_mode(_lib.LSVD_PTR_DEVICE)
    try:
        for i, (m, n) in enumerate(shapes):
            _say("case noalloc", shapes, "call", i, (m, n))
            a = uniform_matrix(0x5EED0A00 + 7 * m + n, m, n)
            u0, s0, vt0, i0 = _solve(h, L.lsvd_small_solve_rectangular_f64, a)
            u1, s1, vt1, i1 = _solve(h, L.lsvd_small_solve_rectangular_f64_noalloc, a)
            k = min(m, n)
            what = f"noalloc {m}x{n}"
            assert_factors(a, u1, s1, vt1, 1e-11, what)
            assert_sigma_close(s1, s0, k, what=what)
            vector_parity(u1, u0, s0, k, what + " U")
            vector_parity(vt1.T, vt0.T, s0, k, what + " V")
            assert i1 == i0, (i1, i0)
            assert np.array_equal(s1, s0) and np.array_equal(u1, u0) and np.array_equal(vt1, vt0), what
    finally:
        h.close()


def test_noalloc_rejects_host_mode_and_oversize():
    L = _lib.lib()
    h = Handle()
    try:
        z = C.c_void_p(16)  # never dereferenced: both calls fail before any launch
        assert L.lsvd_small_solve_rectangular_f64_noalloc(h.raw, z, 8, 8, z, z, z, None) == _lib.LSVD_E_ARG
        h.set_pointer_mode(_lib.LSVD_PTR_DEVICE)
        assert L.lsvd_small_solve_rectangular_f64_noalloc(h.raw, z, 129, 4, z, z, z, None) == _lib.LSVD_E_LIMIT
        assert b"128" in L.lsvd_last_error(h.raw)
        assert L.lsvd_small_solve_rectangular_f64_noalloc(h.raw, z, 0, 4, z, z, z, None) == _lib.LSVD_E_SHAPE
    finally:
        h.close()
