"""The handle's device workspace: entry points that take temporaries from
it allocate nothing on a repeated call of the same size, and give the same
results as the first call (the reused bytes carry no state between calls).
Covers the chip-wide single-matrix route (with the large polar inside it),
Phase-BSS (its batched covariance SVDs and Phase-JADE) and the large
polar on its own."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from lie_cliffalg_analog_svd_amd import LieSvdSmall, newton_schulz_polar, workloads
from lie_cliffalg_analog_svd_amd.bss import separate_batched
from lie_cliffalg_analog_svd_amd.handle import Handle
from tests._util import uniform_matrix
from tests.test_oracle_bss import mixtures

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _twice(h, fn):
    r1 = fn()
    b1, c1 = h.workspace_info()
    r2 = fn()
    b2, c2 = h.workspace_info()
    assert b1 > 0 and (b2, c2) == (b1, c1), ((b1, c1), (b2, c2))
    return r1, r2


def test_large_route_reuses_workspace():
    _say("case workspace large route n=256")
    h = Handle()
    try:
        a = uniform_matrix(0x5EED0B00, 256, 256)
        (u1, s1, vt1), (u2, s2, vt2) = _twice(h, lambda: LieSvdSmall.solve(a, handle=h))
        assert np.array_equal(s1, s2) and np.array_equal(u1, u2) and np.array_equal(vt1, vt2)
    finally:
        h.close()


def test_polar_reuses_workspace():
    _say("case workspace polar 512 x 384")
    h = Handle()
    try:
        a, _, _, _ = workloads.polar_matrix(512, 384, seed=0x5EED0B01, lo=1e-3)
        (q1, ok1), (q2, ok2) = _twice(h, lambda: newton_schulz_polar(a, handle=h))
        assert ok1 and ok2 and np.array_equal(q1, q2)
    finally:
        h.close()


def test_bss_reuses_workspace():
    _say("case workspace bss 16 blocks of 8 x 1024")
    h = Handle()
    try:
        x = np.stack([mixtures(0x5EED0B02 + i, 8, 1024)[1] for i in range(16)])
        r1, r2 = _twice(h, lambda: separate_batched(x, handle=h))
        for a, b in zip(r1, r2):
            assert np.array_equal(a.unmixing, b.unmixing) and np.array_equal(a.separated, b.separated)
    finally:
        h.close()
