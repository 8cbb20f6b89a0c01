#!/usr/bin/env python3
"""Benchmark: batched uneven-size SVD on one MI355X.

Workload (one "step"): one hipGraph launch of the batch plan over
  - 4096 f64 matrices, m ~ U{8..128}, n ~ U{4..m}, A = U diag(s) V^T with
    Haar U, V, s log-spaced to condition 10^k, k ~ U[0, 12], 10% with a
    cluster of 2-4 equal singular values;
  - 8192 square 2x2..4x4 N(0,1) matrices (the LieSvdMicro kernels).
Inputs are resident in HBM when the timed region starts.  value =
matrices/s over the whole batch.

Also reported: the per-launch roofline of the dominant kernel (algorithmic
FP64 flops from the oracle's iteration counts / event-timed launch time),
the sigma error against the imposed spectra, and the CPU oracle (the
restatement of the reference algorithm) timed on all host cores on the
same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and MFMA)
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_flops(m: int, n: int, k_ns: int, sweeps: int) -> float:
    """W = 4 M N^2 [M>N] + 4 k_ns N^3 + 2 N^3 + 6 S N^3 + 2 N^3 for the
    tall-ified M x N problem (QR + compose, NS iterations, Q^T A, Jacobi
    sweeps at 12 N flops per rotation over N(N-1)/2 rotations, Q V)."""
    M, N = max(m, n), min(m, n)
    if N <= 4 and m == n:
        return 0.0
    return 4.0 * M * N * N * (M > N) + (4.0 * k_ns + 4.0 + 6.0 * sweeps) * N ** 3


def polar_leg(h, dev, args):
    """Large NS polar: 4096 x 4096 and 8192 x 2048, A = U diag(s) V^T with Haar
    factors and s log-uniform in [1e-6, 1].  Plain FP64 vs BF16 MFMA
    iterations + FP64 refinement; polar TFLOP/s = iterations x 4 m n^2 / time."""
    import torch

    from lie_cliffalg_analog_svd_amd import LSVD_PREC_BF16, LSVD_PREC_F32, LSVD_PREC_F64, polar, workloads
    res = {}
    for (m, n) in [(4096, 4096), (8192, 2048)]:
        t0 = time.time()
        a, s, u, v = workloads.polar_matrix(m, n)
        exact = u @ v.T
        log(f"polar {m}x{n}: input generated in {time.time() - t0:.1f}s")
        ad = torch.from_numpy(a).to(dev)
        entry = {}
        q64 = None
        for name, prec in [("f64", LSVD_PREC_F64), ("bf16", LSVD_PREC_BF16), ("f32", LSVD_PREC_F32)]:
            qd = torch.empty((m, n), dtype=torch.float64, device=dev)
            info = polar.polar_device(h, ad, m, n, qd, prec)  # warm-up + counts
            torch.cuda.synchronize()
            times = []
            for _ in range(args.polar_reps):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                info = polar.polar_device(h, ad, m, n, qd, prec)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t1)
            sec = min(times)
            iters = info["lowp_iters"] + info["f64_iters"]
            flops = iters * 4.0 * m * n * n
            q = qd.cpu().numpy()
            if name == "f64":
                q64 = q
            entry[name] = {
                "seconds": round(sec, 4), "tflops": round(flops / sec / 1e12, 3),
                "lowp_iters": info["lowp_iters"], "f64_iters": info["f64_iters"], "polar_ok": info["polar_ok"],
                "orth_err_fro": info["orth_err"],
                "err_vs_f64_fro_over_sqrt_n": float(np.linalg.norm(q - q64) / np.sqrt(n)),
                "err_vs_exact_fro_over_sqrt_n": float(np.linalg.norm(q - exact) / np.sqrt(n)),
            }
            log(f"polar {m}x{n} {name}: {entry[name]}")
            del qd
        entry["speedup_bf16_vs_f64"] = round(entry["f64"]["seconds"] / entry["bf16"]["seconds"], 3)
        entry["speedup_f32_vs_f64"] = round(entry["f64"]["seconds"] / entry["f32"]["seconds"], 3)
        res[f"{m}x{n}"] = entry
        del ad
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "bf16"])
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--tiny", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner")
    ap.add_argument("--no-polar", action="store_true", help="skip the large polar leg")
    ap.add_argument("--polar-reps", type=int, default=2)
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the batch shards by matrix with no collective, see DESIGN.md", file=sys.stderr)
        return 2

    import torch

    import lie_cliffalg_analog_svd_amd as L
    from lie_cliffalg_analog_svd_amd import _lib, metrics, workloads
    from lie_cliffalg_analog_svd_amd.handle import BatchPlan, Handle, info_array

    prec = {"f64": L.LSVD_PREC_F64, "f32": L.LSVD_PREC_F32, "bf16": L.LSVD_PREC_BF16}[args.precision]
    t0 = time.time()
    batch = workloads.ragged_batch(args.count, args.tiny)
    mats = batch.mats
    count = len(mats)
    ms = np.array([a.shape[0] for a in mats], dtype=np.int32)
    ns = np.array([a.shape[1] for a in mats], dtype=np.int32)
    log(f"workload: {count} matrices generated in {time.time() - t0:.1f}s")

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)  # default priority
    torch.cuda.set_stream(stream)
    h = Handle(0)
    h.set_pointer_mode(L.LSVD_PTR_DEVICE)
    h.set_stream(stream.cuda_stream)
    if not args.no_tune:
        big = ~((ms == ns) & (ns <= 4))
        t1 = time.time()
        st = _lib.lib().lsvd_autotune_batch(h.raw, int(big.sum()), ms[big].ctypes.data_as(_lib._i32p),
                                            ns[big].ctypes.data_as(_lib._i32p), prec)
        h.check(st)
        log(f"autotune: {time.time() - t1:.1f}s")
    plan = BatchPlan(h, ms, ns, prec)
    lay = plan.layout
    a_pool = torch.from_numpy(workloads.pack(mats)).to(dev)
    u_pool = torch.zeros(lay["u_len"], dtype=torch.float64, device=dev)
    s_pool = torch.zeros(lay["s_len"], dtype=torch.float64, device=dev)
    vt_pool = torch.zeros(lay["vt_len"], dtype=torch.float64, device=dev)
    info_dev = torch.zeros((count, 8), dtype=torch.int32, device=dev)
    log(f"inputs: {a_pool.numel() * 8 / 1e6:.1f} MB in HBM")

    # correctness + per-matrix counts from one run with info
    plan.execute(a_pool, u_pool, s_pool, vt_pool, info_dev)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        plan.execute(a_pool, u_pool, s_pool, vt_pool)
    torch.cuda.synchronize()

    # ---- timed region: exactly K graph launches
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        plan.execute(a_pool, u_pool, s_pool, vt_pool)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_per_step = ev0.elapsed_time(ev1) / args.steps
    value = count / (ms_per_step / 1e3)
    log(f"timed: {args.steps} steps, {ms_per_step:.3f} ms/step (wall {wall / args.steps * 1e3:.3f} ms/step)")

    # ---- per-launch times of each bucket (HIP events on the plan's stream)
    descs, order = plan.describe()
    bucket_ms = plan.time_buckets(a_pool, u_pool, s_pool, vt_pool)
    torch.cuda.synchronize()

    # ---- where the cycles go (per-phase clock64 sums over the batch)
    clocks = torch.zeros((count, 8), dtype=torch.int64, device=dev)
    plan.phase_clocks(a_pool, u_pool, s_pool, vt_pool, clocks)
    ck = clocks.cpu().numpy().sum(axis=0).astype(np.float64)
    names = ["qr", "ns", "p", "jacobi", "u", "repair", "out"]
    phase_split = {nm: round(float(ck[i] / max(ck.sum(), 1.0)), 4) for i, nm in enumerate(names)}
    log(f"phase split: {phase_split}")

    # ---- results to host, sigma errors
    u_h, s_h, vt_h = u_pool.cpu().numpy(), s_pool.cpu().numpy(), vt_pool.cpu().numpy()
    info_h = info_dev.cpu().numpy()
    outs = plan.split(u_h, s_h, vt_h)
    e_rel_max = e_abs_max = 0.0
    worst_recon = worst_orth = 0.0
    for b in range(count):
        u, s, vt = outs[b]
        if batch.sigma_ref[b] is not None:
            e_rel, e_abs = metrics.sigma_errors(s, batch.sigma_ref[b])
            e_rel_max = max(e_rel_max, e_rel)
            e_abs_max = max(e_abs_max, e_abs)
        if b % 37 == 0:
            mm = metrics.compute(mats[b], u, s, vt)
            worst_recon = max(worst_recon, mm.rel_recon)
            worst_orth = max(worst_orth, mm.orth_u, mm.orth_v)

    # ---- CPU baseline: the oracle on all host cores, same workload
    cpu = None
    oracle_info = None
    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        oracle.svd_batch(mats[:64], threads=threads)  # warm-up
        t2 = time.perf_counter()
        ref_out, oracle_info = oracle.svd_batch(mats, threads=threads, with_info=True)
        cpu_s = time.perf_counter() - t2
        cpu = {"value": count / cpu_s, "unit": "matrices/s", "cores": threads, "kind": "port",
               "sample": f"full workload ({count} matrices), C restatement of LieSvdSmall/LieSvdMicro, "
                         f"one matrix per OpenMP thread, {cpu_s:.2f}s"}
        # GPU vs oracle sigma (two-tier) on every matrix
        gvo_rel = gvo_abs = 0.0
        for b in range(count):
            e_rel, e_abs = metrics.sigma_errors(outs[b][1], ref_out[b][1])
            gvo_rel = max(gvo_rel, e_rel)
            gvo_abs = max(gvo_abs, e_abs)
        log(f"cpu oracle: {cpu['value']:.0f} matrices/s on {threads} threads; "
            f"gpu vs oracle sigma rel {gvo_rel:.2e} abs {gvo_abs:.2e}")

    # ---- roofline of the dominant launch
    counts = oracle_info if oracle_info is not None else [
        {"ns_iters": int(info_h[b, 1]), "jacobi_sweeps": int(info_h[b, 2])} for b in range(count)]
    work = np.array([algorithmic_flops(int(ms[b]), int(ns[b]), counts[b]["ns_iters"], counts[b]["jacobi_sweeps"])
                     for b in range(count)])
    dom = int(np.argmax(bucket_ms))
    d = descs[dom]
    items = order[d["first"]:d["first"] + d["count"]] if d["variant"] >= 0 else []
    dom_flops = float(work[items].sum()) if len(items) else 0.0
    achieved = dom_flops / (bucket_ms[dom] / 1e3) / 1e12
    total_tflops = float(work.sum()) / (ms_per_step / 1e3) / 1e12
    traffic = None
    tf_path = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf_path):
        try:
            traffic = json.load(open(tf_path)).get(f"batch_variant{d['variant']}_maxn{d['max_n']}")
        except Exception:
            traffic = None
    roofline = {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": traffic,
                "kernel": f"small_svd_kernel variant {d['variant']} (max N {d['max_n']}, {d['count']} matrices, "
                          f"{d['threads']} threads x {d['grid']} workgroups)",
                "launch_ms": round(bucket_ms[dom], 4)}

    polar_res = None if args.no_polar else polar_leg(h, dev, args)

    out = {
        "metric": "batched SVD matrices/s + polar TFLOP/s, chosen vs plain FP64, at stated sigma err",
        "value": round(value, 1),
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f64": "f64", "f32": "f32+f64", "bf16": "bf16+f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "batched uneven-size SVD: 4096 f64 m~U[8,128], n~U[4,m], kappa 10^U[0,12], "
                               "10% clustered + 8192 square 2..4 (LieSvdMicro), one hipGraph launch",
                   "matrices": count, "precision": args.precision, "parallelism": "single GPU"},
        "sigma_err": {"rel_max_vs_imposed": e_rel_max, "abs_max_vs_imposed": e_abs_max,
                      "rule": "rel over sigma>=1e-3*sigma_max, abs normalised by sigma_max"},
        "residuals_sampled": {"rel_recon_max": worst_recon, "orth_max": worst_orth},
        "fp64_tflops_whole_batch": round(total_tflops, 4),
        "buckets": [{**dd, "ms": round(t, 4)} for dd, t in zip(descs, bucket_ms)],
        "phase_split": phase_split,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "polar": polar_res,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
