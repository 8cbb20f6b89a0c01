"""GPU adjoint (vjp) of the SVD and the polar factor.

Checks, in FP64:
- against the CPU analytic gradient of L = sum_i w_i s_i + <G, U Vt>
  (oracle.svd_loss_grad, closed form), to 1e-10 relative;
- against central finite differences (h = 1e-6) along random directions,
  to 1e-6 relative, on spectra with gaps >= 1e-3 (FD is meaningless when
  the gap is below h);
- a generic cotangent loss L = sum_i beta_i u_i^T P v_i (sign-invariant,
  exercises the guarded 1/(s_j - s_i) part) against central FD;
- finiteness and agreement with the analytic gradient on gap-1e-8 spectra
  (the guard case); forward recomputation vs passing stored factors.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import LieSvdSmall, adjoint, rng, workloads

pytestmark = pytest.mark.gpu


def _gapped(seed, m, n, gap_lo=1e-3, tiny_gap=None):
    u = rng.haar(rng.mix_seed(seed, 1), m, n)
    v = rng.haar(rng.mix_seed(seed, 2), n)
    steps = gap_lo + gap_lo * rng.uniform(rng.mix_seed(seed, 3), n)
    s = 1.0 + np.cumsum(steps[::-1])[::-1] - steps[-1]
    if tiny_gap is not None:
        k = n // 2
        s[k + 1] = s[k] - tiny_gap
    return (u * s[None, :]) @ v.T


def _loss(a, w, g):
    u, s, vt = LieSvdSmall.solve_rectangular(a)
    return float(np.dot(w, s) + np.sum(g * (u @ vt)))


@pytest.mark.parametrize("m,n", [(96, 64), (40, 24), (33, 33)])
def test_loss_grad_vs_cpu_analytic(m, n):
    B = 6
    a = np.stack([_gapped(100 + b, m, n) for b in range(B)])
    w = rng.uniform(7, B * n).reshape(B, n)
    g = rng.normal(8, B * m * n).reshape(B, m, n)
    print("vjp analytic", m, n, flush=True)
    grad = adjoint.svd_loss_grad(a, w, g)
    for b in range(B):
        _, ref = oracle.svd_loss_grad(a[b], w[b], g[b])
        assert np.linalg.norm(grad[b] - ref) <= 1e-10 * np.linalg.norm(ref)


def test_loss_grad_vs_central_fd():
    m, n, h = 96, 64, 1e-6
    for b in range(3):
        a = _gapped(200 + b, m, n)
        w = rng.uniform(300 + b, n)
        g = rng.normal(400 + b, m * n).reshape(m, n)
        grad = adjoint.svd_loss_grad(a, w, g)[0]
        for t in range(3):
            d = rng.normal(500 + 10 * b + t, m * n).reshape(m, n)
            d /= np.linalg.norm(d)
            fd = (_loss(a + h * d, w, g) - _loss(a - h * d, w, g)) / (2 * h)
            an = float(np.sum(grad * d))
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (b, t, fd, an)


def test_generic_cotangents_vs_fd():
    """L = sum_i beta_i u_i^T P v_i: gU = P V diag(beta), gVt = diag(beta) U^T P."""
    m, n, h = 48, 32, 1e-6
    a = _gapped(600, m, n, gap_lo=2e-2)
    p = rng.normal(601, m * n).reshape(m, n)
    beta = 0.5 + rng.uniform(602, n)

    def loss(x):
        u, s, vt = LieSvdSmall.solve_rectangular(x)
        return float(np.sum(beta * np.einsum("ij,ij->j", u, p @ vt.T)))

    u, s, vt = LieSvdSmall.solve_rectangular(a)
    gu = (p @ vt.T) * beta[None, :]
    gvt = beta[:, None] * (u.T @ p)
    grad = adjoint.svd_vjp(a, gu=gu, gvt=gvt, forward=(u[None], s[None], vt[None]))[0]
    grad2 = adjoint.svd_vjp(a, gu=gu, gvt=gvt)[0]  # forward recomputed
    assert np.linalg.norm(grad - grad2) <= 1e-12 * np.linalg.norm(grad)
    for t in range(4):
        d = rng.normal(700 + t, m * n).reshape(m, n)
        d /= np.linalg.norm(d)
        fd = (loss(a + h * d) - loss(a - h * d)) / (2 * h)
        an = float(np.sum(grad * d))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (t, fd, an)


def test_tiny_gap_guard():
    """Gap 1e-8: the loss is gauge invariant, so the guard never enters and
    the result equals the analytic gradient; the generic path stays finite."""
    m, n = 96, 64
    a = _gapped(800, m, n, tiny_gap=1e-8)
    w = rng.uniform(801, n)
    g = rng.normal(802, m * n).reshape(m, n)
    grad = adjoint.svd_loss_grad(a, w, g)[0]
    _, ref = oracle.svd_loss_grad(a, w, g)
    assert np.all(np.isfinite(grad))
    assert np.linalg.norm(grad - ref) <= 1e-8 * np.linalg.norm(ref)
    gu = rng.normal(803, m * n).reshape(m, n)
    gg = adjoint.svd_vjp(a, gu=gu, guard=1e-6)[0]
    assert np.all(np.isfinite(gg)) and np.linalg.norm(gg) < 1e8
