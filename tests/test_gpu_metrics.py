"""Device metrics pass (lsvd_metrics_batched_f64, metrics::compute of
src/lib.rs:149-194) and the full-size property checks it enables.

Pins: exact factors give exactly 0 (diagonal and permutation matrices);
U = I + eps E against the closed forms ||eps E diag(s)||_F / ||s|| and
||eps (E + E^T) + eps^2 E^T E||_F; a planted spectrum perturbation against
its known relative errors (all i and the tail i >= floor(3m/4)); random
thin factors against the host restatement at the FP64 rounding level
(abs 1e-13).  Then the 12288-matrix ragged batch (4096 ragged f64 m in
[8, 128], kappa up to 1e12, plus 8192 squares 2..4): every matrix's
residuals through the device pass (tolerance 1e-11 max(1, n / 32), the
small-kernel parity bar) and its sigma against the imposed spectrum, two-tier
(relative 1e-12 above 1e-3 sigma_max, absolute 64 n eps sigma_max).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from lie_cliffalg_analog_svd_amd import LSVD_PREC_F64_SCALED, LSVD_PTR_DEVICE, metrics, rng, workloads
from lie_cliffalg_analog_svd_amd.handle import BatchPlan, Handle, pack_offsets
from tests._util import EPS

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(h, cases, sigma_refs=None):
    """cases: list of (a, u, s, vt) host arrays -> device metrics (host mode)."""
    ms = np.array([c[0].shape[0] for c in cases], dtype=np.int32)
    ns = np.array([c[0].shape[1] for c in cases], dtype=np.int32)
    lay = pack_offsets(ms, ns)
    pools = [workloads.pack([c[i] for c in cases]) for i in range(4)]
    ref = workloads.pack(sigma_refs) if sigma_refs is not None else None
    return metrics.compute_batched(h, ms, ns, lay, *pools, sigma_ref_pool=ref)


def test_exact_factors_give_zero():
    h = Handle()
    try:
        n = 12
        s = np.linspace(3.0, 0.5, n)
        perm = np.eye(n)[np.argsort(rng.uniform(0x5EED0C00, n))]
        cases = [(np.diag(s), np.eye(n), s, np.eye(n)),
                 (perm * s[None, :], perm, s, np.eye(n))]
        out = _run(h, cases, [s, s])
        for r in out:
            assert r["rel_recon"] == 0.0 and r["orth_u"] == 0.0 and r["orth_v"] == 0.0, r
            assert r["sigma_max_rel"] == 0.0 and r["sigma_tail_rel"] == 0.0, r
    finally:
        h.close()


def test_closed_forms():
    h = Handle()
    try:
        n, eps = 40, 1e-3
        e = rng.normal(0x5EED0C01, n * n).reshape(n, n)
        s = np.linspace(2.0, 1.0, n)
        u = np.eye(n) + eps * e
        want_recon = np.linalg.norm(eps * e * s[None, :]) / np.linalg.norm(s)
        want_orth = np.linalg.norm(eps * (e + e.T) + eps * eps * (e.T @ e))
        # planted spectrum error: got = want (1 + d) with d known, given shuffled
        d = np.abs(rng.normal(0x5EED0C02, n)) * 1e-9
        want_s = np.sort(np.exp(rng.normal(0x5EED0C03, n)))[::-1]
        got_s = want_s * (1.0 + d)
        shuffled = want_s[np.argsort(rng.uniform(0x5EED0C04, n))]
        out = _run(h, [(np.diag(s), u, s, np.eye(n)), (np.diag(got_s), np.eye(n), got_s, np.eye(n))],
                   [s, shuffled])
        r0, r1 = out[0], out[1]
        _say("closed forms", r0, r1)
        assert abs(r0["rel_recon"] - want_recon) <= 1e-12 * want_recon
        assert abs(r0["orth_u"] - want_orth) <= 1e-10 * want_orth
        assert r0["orth_v"] == 0.0
        assert abs(r1["sigma_max_rel"] - d.max()) <= 1e-15 + 1e-6 * d.max()
        tail = (3 * n) // 4
        assert abs(r1["sigma_tail_rel"] - d[tail:].max()) <= 1e-15 + 1e-6 * d[tail:].max()
        assert np.isnan(_run(h, [(np.diag(s), u, s, np.eye(n))])[0]["sigma_max_rel"])
    finally:
        h.close()


@pytest.mark.parametrize("m,n", [(7, 3), (3, 7), (65, 65), (128, 40)])
def test_thin_factors_vs_host(m, n):
    h = Handle()
    try:
        a = rng.normal(0x5EED0C10 + m * n, m * n).reshape(m, n)
        u, s, vt = np.linalg.svd(a, full_matrices=False)
        u = u + 1e-6 * rng.normal(0x5EED0C11, u.size).reshape(u.shape)  # residuals well above rounding
        r = _run(h, [(a, u, s, vt)], [s * (1 + 1e-10)])[0]
        w = metrics.compute(a, u, s, vt, s * (1 + 1e-10))
        for f in ("rel_recon", "orth_u", "orth_v", "sigma_max_rel", "sigma_tail_rel"):
            assert abs(r[f] - getattr(w, f)) <= 1e-13 + 1e-9 * abs(getattr(w, f)), (f, r[f], getattr(w, f))
    finally:
        h.close()


def test_full_batch_every_matrix():
    """Config-2 workload at full size: all 12288 matrices checked."""
    import torch
    _say("case full batch metrics")
    batch = workloads.ragged_batch(4096, 8192)
    mats = batch.mats
    ms = np.array([a.shape[0] for a in mats], dtype=np.int32)
    ns = np.array([a.shape[1] for a in mats], dtype=np.int32)
    dev = torch.device("cuda:0")
    h = Handle()
    h.set_pointer_mode(LSVD_PTR_DEVICE)
    try:
        plan = BatchPlan(h, ms, ns, LSVD_PREC_F64_SCALED)
        lay = plan.layout
        a = torch.from_numpy(workloads.pack(mats)).to(dev)
        u = torch.zeros(lay["u_len"], dtype=torch.float64, device=dev)
        s = torch.zeros(lay["s_len"], dtype=torch.float64, device=dev)
        vt = torch.zeros(lay["vt_len"], dtype=torch.float64, device=dev)
        plan.execute(a, u, s, vt)
        torch.cuda.synchronize()
        out = metrics.compute_batched(h, ms, ns, lay, a, u, s, vt)
        plan.close()
        k = np.minimum(ms, ns)
        tol = 1e-11 * np.maximum(1.0, np.maximum(ms, ns) / 32.0)
        worst = {f: float(np.max(out[f] / tol)) for f in ("rel_recon", "orth_u", "orth_v")}
        _say("worst residual / tolerance", worst)
        assert all(v < 1.0 for v in worst.values()), worst
        # sigma vs imposed, two-tier, on every matrix that has one
        s_h = s.cpu().numpy()
        so = lay["s_off"]
        n_checked = 0
        for b in range(len(mats)):
            ref = batch.sigma_ref[b]
            if ref is None:
                continue
            got = s_h[so[b]:so[b] + k[b]]
            e_rel, e_abs = metrics.sigma_errors(got, ref)
            assert e_rel <= 1e-12 and e_abs <= 64 * k[b] * EPS, (b, ms[b], ns[b], e_rel, e_abs)
            n_checked += 1
        assert n_checked == 4096
        # the device pass agrees with the host restatement on a sample
        u_h, vt_h = u.cpu().numpy(), vt.cpu().numpy()
        uo, vo = lay["u_off"], lay["vt_off"]
        for b in range(0, len(mats), 97):
            mb, nb, kb = int(ms[b]), int(ns[b]), int(k[b])
            w = metrics.compute(mats[b], u_h[uo[b]:uo[b] + mb * kb].reshape(mb, kb), s_h[so[b]:so[b] + kb],
                                vt_h[vo[b]:vo[b] + kb * nb].reshape(kb, nb))
            for f in ("rel_recon", "orth_u", "orth_v"):
                assert abs(out[b][f] - getattr(w, f)) <= 1e-13, (b, f, out[b][f], getattr(w, f))
    finally:
        h.close()
