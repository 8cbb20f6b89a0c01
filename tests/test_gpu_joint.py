"""GPU parity of Phase-JADE (LieSvdJoint::diagonalize_symmetric_with_params)
against the CPU oracle and the reference's own test assertions.

The GPU applies the rotations of a round-robin layer together (they commute);
the oracle applies them one by one.  Parity is therefore in results: the
off-diagonal residual, orthogonality of V, recovery of the planted basis up
to permutation and sign, the joint diagonals (sorted), and the zero-rotation
case -- not in rotation counts.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import (JointDiagonalizationParams, LieSvdJoint, ShapeError, diagonalize_families,
                                         profiles, rng, workloads)

pytestmark = pytest.mark.gpu


def _family(n, k, seed):
    q = profiles.random_orthogonal(n, seed)
    mats = np.stack([q @ np.diag(1.0 + f * 0.3 + np.arange(n) * 0.17) @ q.T for f in range(k)])
    return q, mats


@pytest.mark.parametrize("n,k,seed", [(12, 5, 130), (10, 2, 131), (14, 9, 132), (13, 4, 133), (33, 7, 134)])
def test_joint_family_vs_oracle(n, k, seed):
    """src/lie_svd_joint.rs:548-585 assertions, plus oracle parity."""
    print("joint", n, k, seed, flush=True)
    q, mats = _family(n, k, seed)
    params = JointDiagonalizationParams(40)
    v, diags, tr = LieSvdJoint.diagonalize_symmetric_with_params(list(mats), params)
    vo, do, tro = oracle.joint_diag_sym(mats, oracle.JointParams(40, 1e-12, 1e-14, 6))
    assert tr.final_offdiag < tr.initial_offdiag * 1e-8
    assert tr.rotations > 0
    assert len(diags) == k
    assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-10
    assert abs(tr.initial_offdiag - tro["initial_offdiag"]) < 1e-12 * tro["initial_offdiag"]
    assert workloads.recovery_error(q, v) < 1e-9
    for f in range(k):
        assert np.max(np.abs(np.sort(diags[f]) - np.sort(do[f]))) < 1e-10
        w = v.T @ mats[f] @ v
        assert np.linalg.norm(w - np.diag(np.diag(w))) < 1e-9


def test_joint_already_diagonal_family():
    """src/lie_svd_joint.rs:588-596: no rotation, off <= 1e-14."""
    mats = [np.diag(1.0 + np.arange(8) + k) for k in range(4)]
    v, d, tr = LieSvdJoint.diagonalize_symmetric(mats)
    assert tr.final_offdiag <= 1e-14 and tr.rotations == 0
    assert np.array_equal(v, np.eye(8))
    assert tr.sweeps == 1


def test_joint_trivial_sizes():
    v, d, tr = LieSvdJoint.diagonalize_symmetric([np.array([[3.0]]), np.array([[-2.0]])])
    assert v.shape == (1, 1) and v[0, 0] == 1.0 and d[0][0] == 3.0 and d[1][0] == -2.0
    assert tr.sweeps == 0 and tr.rotations == 0
    with pytest.raises(ShapeError):
        LieSvdJoint.diagonalize_symmetric_with_params([], JointDiagonalizationParams(4))


def test_joint_noisy_families_batched():
    """Several noisy families (the Phase-JADE workload at reduced size) in
    one call: planted basis recovered to the noise level, and the result
    agrees with the oracle run family by family."""
    F, K, n, eps = 3, 8, 32, 1e-6
    fams, qs, ds = workloads.jade_batch(F, K, n, eps, seed=99)
    print("jade batch", fams.shape, flush=True)
    v, d, trs = diagonalize_families(fams)
    for f in range(F):
        vo, do, tro = oracle.joint_diag_sym(fams[f])
        assert np.linalg.norm(v[f].T @ v[f] - np.eye(n)) < 1e-10
        assert trs[f].final_offdiag <= 1.05 * tro["final_offdiag"] + 1e-12 * trs[f].initial_offdiag
        assert workloads.recovery_error(qs[f], v[f]) < 1e-4
        assert abs(workloads.recovery_error(qs[f], v[f]) - workloads.recovery_error(qs[f], vo)) < 1e-6
        for k in range(K):
            assert np.max(np.abs(np.sort(d[f, k]) - np.sort(do[k]))) < 1e-8


def test_joint_single_matrix_is_eigh():
    g = rng.normal_matrix(5, 10, 10)
    p = g + g.T
    v, d, tr = LieSvdJoint.diagonalize_symmetric_with_params([p], JointDiagonalizationParams(60, 1e-14))
    assert np.max(np.abs(np.sort(d[0]) - np.linalg.eigvalsh(p))) < 1e-11
