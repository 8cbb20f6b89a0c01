"""GPU parity of the batched digital polish (solve_with_digital_polish's
finishing step, src/lie_svd_phaseflow.rs:643-658) against the CPU oracle.

Bar: sigma within the two-tier rule of the small solver (relative 1e-12
above 1e-3 sigma_max, 64 n eps sigma_max absolute), the reference's residual
checks on every factor set, and U's singular subspaces matching the oracle's
to the usual cluster bound.  The GEMMs sum in a different order than the
oracle's loops, so agreement is to rounding, not bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import ShapeError, digital_polish, digital_polish_batched
from tests._util import assert_factors, assert_sigma_close, imposed_spectrum, vector_parity

pytestmark = pytest.mark.gpu


def haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))[None, :]


def near_bases(rng, a, eps):
    """Orthonormal bases a few degrees off A's singular bases, as an analog
    or phase-flow preconditioner would hand over."""
    u, _, vt = np.linalg.svd(a)
    n = a.shape[0]
    qu, _ = np.linalg.qr(u + eps * rng.standard_normal((n, n)))
    qv, _ = np.linalg.qr(vt.T + eps * rng.standard_normal((n, n)))
    return qu, qv.T


@pytest.mark.parametrize("sizes,seed", [([1, 2, 3, 4, 5, 8], 500), ([16, 33, 64], 501), ([100, 128, 150], 502)])
def test_polish_batch_vs_oracle(sizes, seed):
    print("polish", sizes, seed, flush=True)
    rng = np.random.default_rng(seed)
    mats, u0s, vt0s, sig = [], [], [], []
    for k, n in enumerate(sizes):
        s = np.sort(np.exp(rng.uniform(np.log(1e-8), 0.0, n)))[::-1]
        a = imposed_spectrum(seed * 10 + k, n, n, s)
        u0, vt0 = near_bases(rng, a, 1e-2) if k % 2 == 0 else (haar(rng, n), haar(rng, n).T)
        mats.append(a); u0s.append(u0); vt0s.append(vt0); sig.append(s)
    out, info = digital_polish_batched(mats, u0s, vt0s, with_info=True)
    for k, (a, (u, s, vt)) in enumerate(zip(mats, out)):
        n = a.shape[0]
        uo, so, vto, io = oracle.polish(a, u0s[k], vt0s[k])
        assert_sigma_close(s, so, n)
        assert_sigma_close(s, sig[k], n)
        assert_factors(a, u, s, vt, 1e-11 * max(n / 32, 1), f"polish n={n}")
        vector_parity(u, uo, so, n, f"polish U n={n}")
        vector_parity(vt.T, vto.T, so, n, f"polish V n={n}")
        assert info[k]["ns_iters"] > 0 or n == 1


def test_polish_identity_bases_and_scaled_schedule():
    """Identity bases: the polish is LieSvdSmall::solve of A itself; the
    scaled step schedule gives the same factors to the bar."""
    print("polish identity", flush=True)
    rng = np.random.default_rng(503)
    mats = [rng.standard_normal((n, n)) for n in (3, 20, 70)]
    eye = [np.eye(a.shape[0]) for a in mats]
    plain = digital_polish_batched(mats, eye, eye)
    scaled = digital_polish_batched(mats, eye, eye, precision=3)
    for a, (u, s, vt), (u2, s2, vt2) in zip(mats, plain, scaled):
        _, so, _, _ = oracle.small_solve(a)
        assert_sigma_close(s, so, a.shape[0])
        assert_sigma_close(s2, so, a.shape[0])
        assert_factors(a, u2, s2, vt2, 1e-11, "scaled polish")


def test_polish_errors():
    with pytest.raises(ShapeError):
        digital_polish(np.zeros((3, 4)), np.eye(3), np.eye(4))
    with pytest.raises(ShapeError):
        digital_polish_batched([np.eye(3)], [np.eye(3)], [])
    u, s, vt = digital_polish(np.diag([3.0, -1.0]), np.eye(2), np.eye(2))
    assert np.allclose(s, [3.0, 1.0]) and np.allclose(u @ np.diag(s) @ vt, np.diag([3.0, -1.0]))
