"""GPU parity of the tensor HO-SVD (LieSvdTensor::hosvd3,
src/lie_svd_tensor.rs:31-62) against the oracle: the mode factors per
cluster of equal Gram eigenvalues, |core| entries, the traces, the
reconstruction, the reference's test and published smoke, batches of
tensors, and the error paths.

The GPU forms each mode Gram in the reference's summation order without FMA
(the oracle's bits); the Gram SVDs then run on the small-SVD kernels, whose
sums run in another order, so factors agree to rounding and are compared up
to sign.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lie_cliffalg_analog_svd_amd import (LieSvdTensor, LsvdError, ShapeError, hosvd3_batched, reconstruct_hosvd3,
                                         rotate_tensor3_batched, tensor_relative_error)
from tests._util import vector_parity
from tests.test_oracle_tensor import smoke_tensor

pytestmark = pytest.mark.gpu


def check_vs_oracle(x, f, what):
    core, u1, u2, u3, tr = oracle.hosvd3(x)
    scale = np.abs(core).max()
    assert np.max(np.abs(np.abs(f.core) - np.abs(core))) < 1e-11 * scale, what
    for m, (ug, uo) in enumerate(zip((f.u1, f.u2, f.u3), (u1, u2, u3))):
        unf = np.moveaxis(x, m, 0).reshape(x.shape[m], -1)
        s = np.linalg.svd(unf @ unf.T, compute_uv=False)
        vector_parity(ug, uo, s, x.shape[m], f"{what} U{m + 1}")
    assert abs(f.trace.initial_offdiag - tr["initial_offdiag"]) <= 1e-13 * max(1.0, tr["initial_offdiag"])
    assert abs(f.trace.final_offdiag - tr["final_offdiag"]) <= 1e-10 * max(1.0, np.sqrt(np.sum(x * x)))
    assert abs(f.trace.superdiag_mass_ratio - tr["superdiag_mass_ratio"]) < 1e-12
    assert tensor_relative_error(x, reconstruct_hosvd3(f)) < 1e-13


def test_reference_unit_test_and_published_smoke():
    print("hosvd smoke", flush=True)
    t5 = smoke_tensor(5, lead=0)
    f = LieSvdTensor.hosvd3(t5)
    assert tensor_relative_error(t5, reconstruct_hosvd3(f)) < 1e-10 and f.trace.superdiag_mass_ratio > 0.95
    t16 = smoke_tensor(16)
    f = LieSvdTensor.hosvd3(t16)
    assert tensor_relative_error(t16, reconstruct_hosvd3(f)) < 1e-14
    assert "%.3e" % f.trace.superdiag_mass_ratio == "9.999e-01" and f.trace.shape == (16, 16, 16)
    check_vs_oracle(t16, f, "smoke16")


@pytest.mark.parametrize("count,shape,seed", [(4, (8, 8, 8), 1), (3, (5, 7, 9), 2), (2, (33, 6, 17), 3),
                                              (1, (64, 4, 4), 4), (6, (1, 3, 2), 5)])
def test_batched_random_tensors_vs_oracle(count, shape, seed):
    print("hosvd batch", count, shape, flush=True)
    rng = np.random.default_rng(seed)
    xs = rng.standard_normal((count,) + shape)
    out = hosvd3_batched(xs)
    assert len(out) == count
    for t in range(count):
        check_vs_oracle(xs[t], out[t], f"{shape}#{t}")


def test_batch_equals_single_calls():
    rng = np.random.default_rng(9)
    xs = rng.standard_normal((5, 6, 7, 8))
    out = hosvd3_batched(xs)
    for t in (0, 4):
        one = LieSvdTensor.hosvd3(xs[t])
        assert np.max(np.abs(one.core - out[t].core)) < 1e-13 * np.abs(one.core).max()
        assert np.max(np.abs(one.u2 - out[t].u2)) < 1e-12


def test_rotate_tensor3_rectangular_vs_einsum():
    rng = np.random.default_rng(11)
    x = rng.standard_normal((3, 4, 5, 6))
    q1, q2, q3 = rng.standard_normal((3, 2, 4)), rng.standard_normal((3, 7, 5)), rng.standard_normal((3, 6, 6))
    out = rotate_tensor3_batched(x, q1, q2, q3)
    want = np.einsum("tai,tbj,tck,tijk->tabc", q1, q2, q3, x)
    assert out.shape == (3, 2, 7, 6) and np.max(np.abs(out - want)) < 1e-13 * np.abs(want).max()
    assert np.max(np.abs(out[1] - oracle.rotate_tensor3(x[1], q1[1], q2[1], q3[1]))) < 1e-13 * np.abs(want).max()


def test_errors():
    with pytest.raises(ShapeError):
        hosvd3_batched(np.zeros((2, 0, 3, 3)))
    with pytest.raises(ShapeError):
        LieSvdTensor.hosvd3(np.zeros((3, 3)))
    with pytest.raises(LsvdError):
        hosvd3_batched(np.zeros((1, 65, 2, 2)))
    assert hosvd3_batched(np.zeros((0, 3, 3, 3))) == []
