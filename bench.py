"""Benchmark of the SPD square-root path on one MI355X.

Workload (one "step"): the batched uneven workload -- 4096 seeded SPD matrices,
n uniform over {16, 32, ..., 256}, cond log-uniform in [1e1, 1e5], even
indices log-uniform spectra and odd indices four tight clusters -- solved to
<= 1e-6 relative Frobenius error against the FP64 root, in one batched solve.

Two paths are timed on the same device-resident batch:
  * plain FP32: FP32 Jacobi to off(A) <= 1e-7 ||A||_F (or its FP32 floor),
    FP32 reconstruction and corrections, binary64 residual;
  * low-bit + refine + learned tier: per matrix, the learned selector
    (features and MLP inference on the GPU, msq/selector.py) picks one of the
    low-bit tiers (BF16 / FP8 / MXFP4 MFMA reconstruction, BF16 corrections,
    Jacobi stopped once the frozen-factor corrections are predicted to
    contract by a tier-specific rate) or the plain path; every matrix that
    misses the binary64 residual check is re-solved on the plain path inside
    the timed region.  --no-selector: the fixed BF16 tier (rate --rate).
``value`` is the low-bit path's matrices/s; every matrix of both paths is
checked against the FP64 root Q diag(sqrt(lambda)) Q^T.  With the selector,
the fixed BF16 tier (rate --rate) is timed as well and reported beside it.

``--workload configs`` runs the single-matrix configurations instead (n = 64
FP32; n = 1024 BF16; n = 4096 with scaled-FP8 and MXFP4 products): error per
correction step against the construction root, solve latency, and at the
n = 4096 shape the codecs' HBM GB/s and the reconstruction product's
TFLOP/s against the MI355X peaks.

``--workload mixture`` runs the learned-tier workload instead (default 2000
matrices, n uniform in [32, 512], spectra one third each log-uniform with
cond up to 1e6, power law k^-alpha with alpha in [0.5, 3], four clusters):
the same paths, with the pass rate (<= 1e-6 against the known root) and the
fallback rate of each; no roofline leg.

``value`` counts the matrices within 1e-6 of their FP64 root (every matrix of
the batch when all pass), per second of the low-bit path.

Also reported:
  * the roofline of the dominant kernel (the Jacobi of orders 129..256,
    tile_jacobi_kernel -- the blocked eigensolver on FP32 MFMA; --jacobi
    scalar runs persistent_jacobi_kernel instead -- on 256 of the batch's
    matrices of orders 241..256, one per CU, with the reference's fixed 8
    sweeps, run on its own and timed with HIP events on the library's
    stream; ``--roofline-only`` runs just these launches, for the profiler).
    Its HBM traffic comes from the rocprofv3 PMC summary in profiles/, used
    only when that summary was taken on the kernel source being run (the
    SHA-256 of kernels_tile_jacobi.hip recorded there);
  * the whole-batch roofline fraction sum_i t_roof,i / t_measured of each
    path, t_roof,i = 9 S n^2 (n - 1) / P_fp32 + (2 + 8 k) n^3 / P_tier +
    2 k n^3 / P_fp64 with the sweeps S and corrections k each matrix ran;
  * a control: FP32 products with the low-bit tier's Jacobi stop rule, which
    separates the stop rule's gain from the low-bit products';
  * a CPU baseline: the reference's CPU plain path (1_baseline/spd.cpp:348-430,
    LAPACK ssyevd + BLAS, restated in C++ in oracle/cpu_baseline.cpp) over the
    whole batch, one matrix per host core with single-threaded BLAS, timed in
    whole passes for about --cpu-budget seconds.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "mixed-mlx_amd")]

METRIC = "SPD sqrts/sec at ≤1e-6 rel err vs FP64: low-bit+refine vs plain FP32, 1 GPU"
FP32_PEAK_TFLOPS = 157.3
# HBM traffic of the roofline kernel: profiles/tile_jacobi_n256_hbm_traffic_pmc.md
# records the bytes per launch and the SHA-256 of the kernel source it was
# measured on.
TRAFFIC_PROFILE = ROOT / "profiles" / "tile_jacobi_n256_hbm_traffic_pmc.md"
JACOBI_SOURCE = ROOT / "mixed-mlx_amd" / "csrc" / "kernels_tile_jacobi.hip"
# MI355X peaks (TFLOP/s): FP32 vector and matrix, FP64, and the dense MFMA
# rates of the product formats.
P_FP32, P_FP64 = 157.3, 78.6
TIER_PEAKS = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0, "mxfp4": 10100.0}


def measured_traffic():
    """(bytes per launch, source hash) from the PMC summary when it was taken
    on the kernel source of this tree, else (None, hash)."""
    import hashlib
    digest = hashlib.sha256(JACOBI_SOURCE.read_bytes()).hexdigest()
    try:
        fields = dict(line.split(":", 1) for line in TRAFFIC_PROFILE.read_text().splitlines()
                      if line.startswith(("source_sha256:", "traffic_bytes_per_launch:")))
    except (OSError, ValueError):
        return None, digest
    if fields.get("source_sha256", "").strip() != digest:
        return None, digest
    return float(fields["traffic_bytes_per_launch"]), digest


def batch_roofline(orders, sweeps, corrections, tiers, seconds):
    """sum_i t_roof,i / t_measured: per matrix, the Jacobi at the FP32 peak,
    reconstruction and corrections at the tier's product peak, the binary64
    residuals at the FP64 peak."""
    t = 0.0
    for n, S, k, tier in zip(orders, sweeps, corrections, tiers):
        n = float(n)
        t += (9.0 * S * n * n * (n - 1) / (P_FP32 * 1e12)
              + (2.0 + 8.0 * k) * n ** 3 / (TIER_PEAKS[tier] * 1e12)
              + 2.0 * k * n ** 3 / (P_FP64 * 1e12))
    return {"sum_t_roof_s": t, "t_measured_s": seconds, "frac": t / seconds}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def jacobi_flops(orders, sweeps):
    """9 S n^2 (n - 1): 18 n flops per rotation, n (n - 1) / 2 rotations a sweep."""
    return sum(9.0 * s * n * n * (n - 1) for n, s in zip(orders, sweeps))


ROOFLINE_MATRICES = 256  # one matrix per CU
ROOFLINE_SWEEPS = 8      # the reference's fixed sweep count (src/metal_solver.cpp:210-213)


def roofline_leg(args, mats, orders, rt, stream, torch, np, capi, dev):
    """Dominant kernel: the Jacobi of orders 129..256 (tile_jacobi_kernel, the
    blocked eigensolver, or persistent_jacobi_kernel with --jacobi scalar: one
    workgroup per CU, the whole solve of a matrix in LDS) on 256 of the
    batch's matrices of orders 241..256 -- one per CU, so every workgroup takes
    one job -- with the reference's fixed 8 sweeps, so that every workgroup
    does the same work and the launch time is the kernel's time per matrix.
    In the batched solve the same kernel takes the jobs of every order 129..256
    and stops per matrix."""
    big = [i for i, n in enumerate(orders) if n > 240][:ROOFLINE_MATRICES]
    sub_orders = [orders[i] for i in big]
    _, sub_flat = capi.pack_batch([mats[i]["A"] for i in big])
    _, _, _, stotal, svtotal = capi.batch_layout(sub_orders)
    SA = torch.from_numpy(sub_flat).to(dev)
    SV = torch.zeros(stotal, dtype=torch.float32, device=dev)
    SD = torch.zeros(svtotal, dtype=torch.float32, device=dev)
    SR = torch.zeros(svtotal, dtype=torch.float32, device=dev)
    jopt = capi.jacobi_options(fixed_sweeps=ROOFLINE_SWEEPS)
    reps = rt.jacobi_batched(sub_orders, SA, SV, SD, SR, jopt)  # warm
    durations = []
    for _ in range(args.roofline_reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        reps = rt.jacobi_batched(sub_orders, SA, SV, SD, SR, jopt)
        e1.record(stream)
        stream.synchronize()
        durations.append(e0.elapsed_time(e1) * 1e-3)
    jt = float(np.mean(durations))
    sweeps = [r.sweeps for r in reps]
    flops = jacobi_flops(sub_orders, sweeps)
    achieved = flops / jt / 1e12
    traffic, digest = measured_traffic()
    blocked = args.jacobi == "blocked"
    return {"bound": "fp32 mfma" if blocked else "fp32 valu", "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
            "traffic": traffic if len(mats) == 4096 and blocked else None,
            "kernel_source_sha256": digest,
            "kernel": (f"{'tile_jacobi_kernel' if blocked else 'persistent_jacobi_kernel'} "
                       f"({len(big)} matrices of orders 241..256, "
                       f"one per CU, {ROOFLINE_SWEEPS} fixed sweeps; "
                       f"{flops / 1e12:.3f} TFLOP per launch at 9 S n^2 (n - 1), "
                       f"{jt * 1e3:.2f} ms mean of {len(durations)} launches)")}


def cpu_baseline(mats, budget_s):
    """Whole passes over the batch with one matrix per host thread until about
    budget_s seconds are spent (at least one pass); matrices/s over all passes,
    and the last pass checked against the known roots."""
    import numpy as np

    from oracle import cpu_baseline as CB

    threads = CB.host_threads()
    batch = CB.Batch([m["A"] for m in mats])
    passes, total = 0, 0.0
    while passes == 0 or total < budget_s:
        secs, corr, res = CB.solve_batch(batch, threads, max_corrections=12, tolerance=1e-12)
        passes += 1
        total += secs
    errs = np.array([np.linalg.norm(batch.root(i) - m["X"]) / np.linalg.norm(m["X"])
                     for i, m in enumerate(mats)])
    return {"value": round(passes * len(mats) / total, 2), "unit": "matrices/s",
            "cores": threads, "kind": "port",
            "sample": (f"the whole batch ({len(mats)} matrices) x {passes} passes in "
                       f"{total:.1f} s: the reference CPU plain path (ssyevd, sgemm, dgemm "
                       f"residual, FP32 corrections to residual 1e-12, at most 12) in C++, one "
                       f"matrix per thread, {threads} threads, single-threaded BLAS"),
            "within_1e-6": int((errs <= 1e-6).sum()), "max_rel_error": float(errs.max()),
            "mean_corrections": float(corr.mean()), "failed": int((corr < 0).sum()),
            "nproc": os.cpu_count(), "cpu_model": CB.cpu_model(), "blas": CB.blas_config()}


# MI355X dense peaks (TFLOP/s) of the product types, and HBM bandwidth.
PRODUCT_PEAKS = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0, "mxfp4": 10100.0}
HBM_PEAK_GBS = 8000.0


def configs_leg(rt, torch, np, capi, G, MS, cpu=True):
    """The single-matrix configurations, one JSON record each: per-step
    relative Frobenius error against the construction root (a solve stopped
    after k corrections, k = 0..K: the Jacobi and the products are
    deterministic, so the k-th solve is the k-th step of the K-step one),
    the solve's latency, and for n = 4096 the codecs' HBM bandwidth and the
    products' MFMA throughput."""
    from oracle import solver as S

    def event_time(fn, reps=5):
        stream = torch.cuda.current_stream()
        fn()
        stream.synchronize()
        out = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            stream.synchronize()
            out.append(e0.elapsed_time(e1) * 1e-3)
        return float(np.median(out))

    def steps(A, Xc, basis, corr, K, jacobi):
        o, aflat = capi.pack_batch([A])
        _, _, _, total, _ = capi.batch_layout(o)
        Ad = torch.from_numpy(aflat).cuda()
        Xd = torch.zeros(total, dtype=torch.float64, device="cuda")
        errs, rep = [], None
        for k in range(K + 1):
            opts = capi.solve_options(basis, corr, jacobi, min_corrections=k, max_corrections=k,
                                      residual_tolerance=0.0)
            rep = rt.solve_batched(o, Ad, Xd, opts)[0]
            errs.append(float(S.relative_error(capi.unpack_batch(o, Xd.cpu().numpy())[0], Xc)))
        opts = capi.solve_options(basis, corr, jacobi, min_corrections=K, max_corrections=K,
                                  residual_tolerance=0.0)
        t = event_time(lambda: rt.solve_batched(o, Ad, Xd, opts), reps=3)
        return {"error_per_step": errs, "sweeps": rep["sweeps"],
                "residual_per_step": rep["residuals"], "seconds": round(t, 5)}

    jac = capi.jacobi_options(off_tol=1e-7)
    records = []
    # n = 64, cond 1e2, FP32 end to end, one correction.
    A, Xc, _ = G.single_matrix(0, 64, 1e-2)
    r = {"config": "n=64 cond 1e2 fp32, 1 correction",
         **steps(A, Xc, capi.FP32, capi.FP32, 1, jac)}
    if cpu:
        import time as _t
        t0 = _t.perf_counter()
        for _ in range(20):
            S.accelerate_mixed(A, None, max_corrections=1, tolerance=0.0)
        r["cpu_reference_seconds"] = round((_t.perf_counter() - t0) / 20, 6)
    records.append(r)
    # n = 1024, cond 1e4, BF16 storage and products, FP64
    # residual, two corrections.
    A, Xc, _ = G.single_matrix(1, 1024, 1e-4)
    r = {"config": "n=1024 cond 1e4 bf16, 2 corrections",
         **steps(A, Xc, capi.BF16, capi.BF16, 2, jac)}
    if cpu:
        import time as _t
        t0 = _t.perf_counter()
        S.accelerate_mixed(A, None, max_corrections=2, tolerance=0.0)
        r["cpu_reference_seconds"] = round(_t.perf_counter() - t0, 4)
    records.append(r)
    # n = 4096, cond 1e3, products in scaled FP8 E4M3 and in
    # MXFP4, up to three corrections: every product in the tier, and the
    # tier's reconstruction with BF16 corrections.
    A, Xc, _ = G.single_matrix(2, 4096, 1e-3)
    for name, p in (("fp8", capi.FP8_E4M3FN_SCALED), ("mxfp4", capi.MXFP4_E2M1)):
        records.append({"config": f"n=4096 cond 1e3 {name} products, 3 corrections",
                        **steps(A, Xc, p, p, 3, jac)})
        records.append({"config": f"n=4096 cond 1e3 {name} reconstruction, bf16 corrections",
                        **steps(A, Xc, p, capi.BF16, 3, jac)})
    # Codecs at the n = 4096 shape: HBM bytes 4N (float) + N bits / 8 (+ N / 32
    # scale bytes for MX) per direction, over the time of one call.  The codec
    # entry points wait for their kernels, so a call's time is the kernel's
    # (about 15 us at this size) plus the host's launch and wait; the kernels'
    # own rates come from a kernel trace.
    n = 4096
    N = n * n
    g = torch.Generator(device="cuda").manual_seed(7)
    src = torch.randn(N, generator=g, device="cuda", dtype=torch.float32) / n ** 0.5
    dst = torch.empty_like(src)
    codecs = {}
    for name, p, bits in (("bf16", capi.BF16, 16), ("fp8_e4m3fn", capi.FP8_E4M3FN, 8),
                          ("mxfp4", capi.MXFP4_E2M1, 4)):
        vals = torch.empty(capi.storage_bytes(p, N), dtype=torch.uint8, device="cuda")
        scl = torch.empty(max(1, capi.scale_bytes(p, N)), dtype=torch.uint8, device="cuda")
        nbytes = 4 * N + N * bits // 8 + capi.scale_bytes(p, N)
        te = event_time(lambda: rt.quantize(p, src, vals, scl))
        td = event_time(lambda: rt.decode(p, vals, scl, dst))
        codecs[name] = {"encode_gbs": round(nbytes / te / 1e9, 1),
                        "decode_gbs": round(nbytes / td / 1e9, 1),
                        "bytes": nbytes}
    vals = torch.empty(N, dtype=torch.uint8, device="cuda")
    sig = torch.empty(1, dtype=torch.float32, device="cuda")
    te = event_time(lambda: rt.quantize_fp8_scaled(src, vals, sig))
    codecs["fp8_e4m3fn_scaled"] = {"encode_gbs": round((5 * N) / te / 1e9, 1), "bytes": 5 * N}
    # Products at n = 4096: the reconstruction X0 = G G^T (2 n^3 flops), the
    # call's row scaling and operand encode included.
    o = [n]
    _, _, _, total, vtotal = capi.batch_layout(o)
    # A random operand with the basis' scale (entries ~ 1 / sqrt(n)): the
    # product's time does not depend on the operand being orthogonal.
    V = (torch.randn(n * n, generator=g, device="cuda", dtype=torch.float32) / n ** 0.5)
    roots = torch.rand(vtotal, generator=g, device="cuda", dtype=torch.float32) + 0.5
    out = torch.empty(total, dtype=torch.float32, device="cuda")
    products = {}
    for name, p in (("fp32", capi.FP32), ("bf16", capi.BF16), ("fp8", capi.FP8_E4M3FN_SCALED),
                    ("mxfp4", capi.MXFP4_E2M1)):
        t = event_time(lambda: rt.reconstruct(p, o, V, roots, out))
        tf = 2.0 * n ** 3 / t / 1e12
        products[name] = {"seconds": round(t, 6), "tflops": round(tf, 1),
                          "frac_of_peak": round(tf / PRODUCT_PEAKS[name], 4)}
    return {"configs": records, "codecs_n4096": codecs, "hbm_peak_gbs": HBM_PEAK_GBS,
            "reconstruction_n4096": products}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=None,
                    help="matrices (default 4096; mixture workload 2000)")
    ap.add_argument("--workload", choices=["uneven", "mixture", "configs"], default="uneven")
    ap.add_argument("--tier", default="bf16")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rate", type=float, default=0.5)
    ap.add_argument("--roofline-only", action="store_true")
    ap.add_argument("--profile-path", choices=["plain", "fixed", "learned"], default=None,
                    help="time only this path (warm-up and steps), for the profiler")
    ap.add_argument("--roofline-reps", type=int, default=5)
    ap.add_argument("--no-selector", action="store_true",
                    help="fixed tier instead of the learned tier selector")
    ap.add_argument("--no-graphs", action="store_true",
                    help="eager launches instead of per-batch HIP graph replay")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--jacobi", choices=["blocked", "scalar"], default="blocked",
                    help="eigensolver of the orders up to 256: blocked (8-index blocks, "
                         "FP32 MFMA products) or scalar (one rotation per pair)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the batched SPD solve shards by matrix with no "
              "exchange, so N GPUs is N independent copies of this run; --gpus must be 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from msq import capi, generators as G, selector as SEL, solver as MS

    if args.workload == "configs":
        torch.cuda.set_device(0)
        rt = capi.Runtime(0)
        rec = configs_leg(rt, torch, np, capi, G, MS, cpu=not args.no_cpu)
        for r in rec["configs"]:
            log(r)
        log("codecs", rec["codecs_n4096"])
        log("reconstruction", rec["reconstruction_n4096"])
        print(json.dumps({"metric": "single-matrix configurations: error per step, latency, "
                                    "codec GB/s and product TFLOP/s at n = 4096",
                          **rec}), flush=True)
        rt.close()
        return

    t = time.perf_counter()
    if args.workload == "uneven":
        mats = G.batched_uneven(count=args.count or 4096)
        workload = ("batched uneven: 4096 SPD matrices, n in {16,...,256}, cond log-uniform "
                    "in [1e1, 1e5], half clustered spectra")
    else:
        mats = []
        for m in G.mixture(count=args.count or 2000, config=4, order_range=(32, 512)):
            mats.append(m)
            if len(mats) % 2000 == 0:
                log(f"generated {len(mats)} matrices")
        workload = (f"learned tier selection: {len(mats)} SPD matrices, n uniform in [32, 512], "
                    "spectra log-uniform (cond <= 1e6) / power law k^-alpha (alpha in "
                    "[0.5, 3]) / 4 clusters")
    orders = [m["order"] for m in mats]
    _, aflat = capi.pack_batch([m["A"] for m in mats])
    lds, offs, voffs, total, vtotal = capi.batch_layout(orders)
    # The orders as one uint32 array for the timed calls (a list is converted
    # again by every call).
    batch_orders = np.ascontiguousarray(orders, dtype=np.uint32)
    log(f"generated {len(mats)} matrices in {time.perf_counter() - t:.1f} s")

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    rt = capi.Runtime(0, stream=stream)
    rt.set_jacobi_method(capi.JACOBI_BLOCKED_MFMA if args.jacobi == "blocked"
                         else capi.JACOBI_SCALAR)
    tile_table = None
    if not args.no_autotune:
        t = time.perf_counter()
        tile_table = rt.autotune(256)
        log(f"autotuned GEMM tiles in {time.perf_counter() - t:.1f} s")
    if not args.no_graphs:
        rt.set_graphs(True)
    A = torch.from_numpy(aflat).to(dev)
    X = torch.zeros(total, dtype=torch.float64, device=dev)

    def check_all(Xdev, label):
        """(worst relative error, matrices within 1e-6) against the known roots."""
        Xh = Xdev.cpu().numpy()
        worst, passed = 0.0, 0
        for m, off, ld in zip(mats, offs, lds):
            n = m["order"]
            Xi = Xh[off:off + ld * ld].reshape(ld, ld, order="F")[:n, :n]
            Xr = m["X"]
            err = float(np.linalg.norm(Xi - Xr) / np.linalg.norm(Xr))
            worst = max(worst, err)
            passed += err <= 1e-6
        log(f"{label}: worst relative error vs FP64 root {worst:.3e}, "
            f"{passed}/{len(mats)} within 1e-6")
        return worst, passed

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        stream.synchronize()
        t0 = time.perf_counter()
        out = None
        for _ in range(args.steps):
            out = fn()
        stream.synchronize()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps, out

    plain = MS.plain_fp32_options()
    if args.roofline_only:
        log("roofline", roofline_leg(args, mats, orders, rt, stream, torch, np, capi, dev))
        rt.close()
        return
    if args.profile_path:
        if args.profile_path == "plain":
            fn = lambda: rt.solve_batched(batch_orders, A, X, plain)  # noqa: E731
        elif args.profile_path == "fixed":
            opts = MS.tier_options(args.tier, rate=args.rate, max_corrections=12)
            fn = lambda: MS.solve_with_fallback(rt, batch_orders, A, X, opts)  # noqa: E731
        else:
            sel = SEL.Selector.load()
            fn = lambda: SEL.solve_selected(rt, sel, batch_orders, A, X)  # noqa: E731
        dt, _ = timed(fn)
        log(f"{args.profile_path}: {dt * 1e3:.1f} ms/step")
        rt.close()
        return
    dt_plain, rep_plain = timed(lambda: rt.solve_batched(batch_orders, A, X, plain))
    err_plain, pass_plain = check_all(X, "plain fp32")
    ok_plain = int(rep_plain.converged.sum())
    log(f"plain fp32: {dt_plain * 1e3:.1f} ms/step, {len(mats) / dt_plain:.0f} matrices/s, "
        f"converged {ok_plain}/{len(mats)}, mean sweeps "
        f"{np.mean(rep_plain.sweeps):.2f}, mean corrections "
        f"{np.mean(rep_plain.corrections):.2f}")

    use_selector = not args.no_selector and SEL.WEIGHTS.exists()
    if use_selector:
        sel = SEL.Selector.load()
        chosen = []

        def low_bit():
            reports, fb, tiers = SEL.solve_selected(rt, sel, batch_orders, A, X)
            chosen[:] = [tiers]
            return reports, fb
    tier = MS.tier_options(args.tier, rate=args.rate, max_corrections=12)

    def fixed_tier():
        return MS.solve_with_fallback(rt, batch_orders, A, X, tier)
    if not use_selector:
        low_bit = fixed_tier
    label = "learned tier" if use_selector else f"fixed {args.tier} tier, rate {args.rate}"
    dt_low, (rep_low, fallback) = timed(low_bit)
    err_low, pass_low = check_all(X, label)
    log(f"{label}: {dt_low * 1e3:.1f} ms/step, {len(mats) / dt_low:.0f} "
        f"matrices/s, fallback {len(fallback)}/{len(mats)}, mean sweeps "
        f"{np.mean(rep_low.sweeps):.2f}")
    fixed = None
    if use_selector:
        # The method without the learned part, on the same batch.
        dt_fix, (rep_fix, fb_fix) = timed(fixed_tier)
        err_fix, pass_fix = check_all(X, f"fixed {args.tier} tier")
        fixed = {"tier": args.tier, "correction_rate": args.rate,
                 "matrices_per_s": round(pass_fix / dt_fix, 2),
                 "ms_per_step": round(dt_fix * 1e3, 3), "max_rel_error": err_fix,
                 "within_1e-6": pass_fix, "fallbacks": len(fb_fix),
                 "mean_sweeps": float(np.mean(rep_fix.sweeps))}
        log(f"fixed {args.tier} tier: {dt_fix * 1e3:.1f} ms/step, "
            f"{len(mats) / dt_fix:.0f} matrices/s, fallback {len(fb_fix)}/{len(mats)}")
    # Control: FP32 products with the same Jacobi stop rule as the fixed tier,
    # so the low-bit products' share of the gain is measured on its own.
    control_opts = MS.tier_options("fp32", rate=args.rate, max_corrections=12)
    dt_ctl, (rep_ctl, fb_ctl) = timed(
        lambda: MS.solve_with_fallback(rt, batch_orders, A, X, control_opts))
    err_ctl, pass_ctl = check_all(X, f"fp32 products, rate {args.rate}")
    control = {"products": "fp32", "correction_rate": args.rate,
               "matrices_per_s": round(pass_ctl / dt_ctl, 2),
               "ms_per_step": round(dt_ctl * 1e3, 3), "max_rel_error": err_ctl,
               "within_1e-6": pass_ctl, "fallbacks": len(fb_ctl),
               "mean_sweeps": float(np.mean(rep_ctl.sweeps)),
               "mean_corrections": float(np.mean(rep_ctl.corrections))}
    log(f"fp32 products, rate {args.rate}: {dt_ctl * 1e3:.1f} ms/step")

    # Whole-batch roofline fraction of each path with the S and k each matrix
    # ran (a matrix re-solved on the plain path counts with that solve).
    if use_selector:
        low_tiers = [SEL.choice_format(int(k)) for k in chosen[0]]
    else:
        low_tiers = [args.tier] * len(mats)
    for i in fallback:
        low_tiers[i] = "fp32"
    batch_roof = {
        "plain_fp32": batch_roofline(orders, rep_plain.sweeps, rep_plain.corrections,
                                     ["fp32"] * len(mats), dt_plain),
        "low_bit": batch_roofline(orders, rep_low.sweeps, rep_low.corrections, low_tiers,
                                  dt_low),
    }
    log("batch roofline", batch_roof)

    roofline = None
    if args.workload == "uneven":
        roofline = roofline_leg(args, mats, orders, rt, stream, torch, np, capi, dev)
        log("roofline", roofline)

    cpu = None if args.no_cpu else cpu_baseline(mats, args.cpu_budget)
    if cpu:
        log("cpu baseline", cpu)

    line = {
        "metric": METRIC,
        "value": round(pass_low / dt_low, 2),
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_low * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 jacobi / bf16 mfma / fp64 residual",
        "data": "synthetic (seeded Haar-random SPD matrices)",
        "config": {"workload": workload,
                   "count": len(mats), "tier": args.tier, "correction_rate": args.rate,
                   "hip_graphs": not args.no_graphs,
                   "tile_table": tile_table.tolist() if tile_table is not None else None},
        "plain_fp32": {"matrices_per_s": round(pass_plain / dt_plain, 2),
                       "ms_per_step": round(dt_plain * 1e3, 3),
                       "max_rel_error": err_plain, "converged": ok_plain,
                       "within_1e-6": pass_plain},
        "low_bit": {"matrices_per_s": round(pass_low / dt_low, 2),
                    "all_matrices_per_s": round(len(mats) / dt_low, 2),
                    "selector": ({"tier_counts": dict(zip(
                        SEL.CHOICES,
                        np.bincount(chosen[0], minlength=len(SEL.CHOICES)).tolist())),
                        # A batch run on the fixed tier as a whole stops its
                        # large orders at the looser rate (batch_options).
                        "fixed_batch_large_orders_at_rate": (
                            {str(SEL.FIXED_BATCH_LARGE_RATE): int(np.sum(
                                np.asarray(orders) > SEL.FIXED_BATCH_LARGE_ORDER))}
                            if np.all(chosen[0] == SEL.FIXED) else None)}
                                 if use_selector else None),
                    "max_rel_error": err_low, "within_1e-6": pass_low,
                    "fallbacks": len(fallback),
                    "mean_sweeps": float(np.mean(rep_low.sweeps)),
                    "mean_corrections": float(np.mean(rep_low.corrections))},
        "speedup_vs_plain_fp32": round((pass_low / dt_low) / max(pass_plain / dt_plain, 1e-30), 3),
        "fixed_tier": fixed,
        "fp32_products_same_stop": control,
        "batch_roofline": batch_roof,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    rt.close()


if __name__ == "__main__":
    main()
