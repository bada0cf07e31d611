"""The batched workloads at their own shapes, through the C ABI.

* The batched uneven workload whole: 4096 matrices, n in {16, ..., 256},
  cond 10^U[1, 5], half clustered spectra, solved in graph mode (eager,
  captured, replayed) with the BF16 tier -- every matrix within 1e-6 of the
  construction root, and the replay equal to the eager solve bit for bit.
* The learned-tier workload on a seeded sample: n in [32, 512] (orders above
  256 included), log-uniform (cond up to 1e6), power-law and clustered
  spectra, through the learned selector with the plain-path fallback (second
  Jacobi pass for unresolved spectra): every matrix within 1e-6.
"""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, say
from msq import capi
from msq import generators as G
from msq import selector as SEL
from msq import solver as MS

pytestmark = pytest.mark.gpu


def errors(mats, orders, Xflat):
    lds, offs, _, _, _ = capi.batch_layout(orders)
    out = np.empty(len(mats))
    for i, (m, off, ld) in enumerate(zip(mats, offs, lds)):
        n = m["order"]
        Xi = Xflat[off:off + ld * ld].reshape(ld, ld, order="F")[:n, :n]
        out[i] = np.linalg.norm(Xi - m["X"]) / np.linalg.norm(m["X"])
    return out


def test_batched_uneven_full_batch_in_graph_mode(runtime):
    mats = G.batched_uneven(count=4096)
    orders = [m["order"] for m in mats]
    o, aflat = capi.pack_batch([m["A"] for m in mats])
    _, _, _, total, _ = capi.batch_layout(o)
    A = dev(aflat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    opts = MS.tier_options("bf16")
    runtime.set_graphs(True)
    try:
        results = []
        for call in range(3):  # eager, captured, replayed
            X.zero_()
            reps = runtime.solve_batched(o, A, X, opts)
            results.append((host(X).copy(), reps.status.copy(), reps.corrections.copy()))
            say(f"uneven 4096 call {call}: converged {int(reps.converged.sum())}")
    finally:
        runtime.set_graphs(False)
    for got, status, corr in results[1:]:
        assert np.array_equal(got, results[0][0])
        assert np.array_equal(status, results[0][1]) and np.array_equal(corr, results[0][2])
    err = errors(mats, orders, results[0][0])
    say(f"uneven 4096: worst error {err.max():.3e}")
    assert (results[0][1] == capi.CONVERGED).all()
    assert err.max() <= 1e-6


def test_learned_tier_workload_sample_with_fallback(runtime):
    mats = list(G.mixture(count=96, config=4, order_range=(32, 512)))
    kinds = {m["kind"] for m in mats}
    orders = [m["order"] for m in mats]
    assert kinds == {"log_uniform", "power_law", "clustered"}
    assert max(orders) > 256 and min(orders) < 128
    o, aflat = capi.pack_batch([m["A"] for m in mats])
    _, _, _, total, _ = capi.batch_layout(o)
    A = dev(aflat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    sel = SEL.Selector.load()
    reports, fallback, tiers = SEL.solve_selected(runtime, sel, o, A, X)
    err = errors(mats, orders, host(X))
    conv = reports.converged
    say(f"learned sample: {int(conv.sum())}/{len(mats)} converged, {len(fallback)} fallbacks, "
        f"tiers {np.bincount(tiers, minlength=len(SEL.TIERS)).tolist()}, "
        f"worst error {err.max():.3e}, refined "
        f"{sum(int(reports[i]['refined']) for i in range(len(mats)))}")
    for i in range(len(mats)):
        if conv[i]:
            assert err[i] <= 1e-6, (i, mats[i]["kind"], mats[i]["condition"], err[i])
    assert conv.all(), [(i, reports[i]["status"], mats[i]["condition"])
                        for i in np.flatnonzero(~conv)]
