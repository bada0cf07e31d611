"""Measures the latency of the plain-path fallback solve per matrix order and
adds it to the selector's weights file (msq/data/selector.npz).

    python mixed-mlx_amd/msq/fallback_latency.py

A matrix whose tier misses the binary64 residual check is solved again on
the plain path in a second batched call, whose time is that of its slowest
matrix rather than the matrix's share of a batch.  This times that call for
one matrix of each order of a grid (the median of five calls after a first,
on seeded log-uniform spectra with cond 1e4), which the selector
interpolates in n^3 as the cost of a failed choice.
"""
from __future__ import annotations

import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from msq import capi, generators as G, selector as SEL, solver as MS  # noqa: E402

ORDERS = (16, 32, 64, 96, 128, 160, 192, 224, 256, 320, 384, 448, 512)


def measure(rt, orders=ORDERS):
    import torch

    out = []
    for k, n in enumerate(orders):
        A, _, _ = G.single_matrix(20, n, 1e-4, index=k)
        o, flat = capi.pack_batch([A])
        _, _, _, total, _ = capi.batch_layout(o)
        Ad = torch.from_numpy(flat).cuda()
        X = torch.zeros(total, dtype=torch.float64, device="cuda")
        opts = MS.fallback_options()
        rt.solve_batched(o, Ad, X, opts)
        times = []
        for _ in range(5):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rt.solve_batched(o, Ad, X, opts)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        out.append(float(np.median(times)))
        print(f"fallback solve, n = {n}: {out[-1]:.2f} ms", flush=True)
    return np.array(orders), np.array(out)


def main():
    import torch

    torch.cuda.set_device(0)
    rt = capi.Runtime(0)
    orders, ms = measure(rt)
    with np.load(SEL.WEIGHTS, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    data["fallback_orders"] = orders
    data["fallback_ms"] = ms
    np.savez(SEL.WEIGHTS, **data)
    print(f"saved {SEL.WEIGHTS}", flush=True)
    rt.close()


if __name__ == "__main__":
    main()
