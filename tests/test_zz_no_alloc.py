"""The allocation-free Jacobi entry point (msq_jacobi_batched_noalloc) against
the allocating one (msq_jacobi_batched), through the C ABI: the same
eigenvectors, diagonals, roots and reports bit for bit, on two calls in a row
(the second with a smaller batch, so stale scratch entries of the first call
would show), plus the host-side rejections."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, say
from msq import capi
from msq import generators as G

pytestmark = pytest.mark.gpu


def batch(orders, seed):
    mats = [G.single_matrix(7, n, 10.0 ** -(1 + (i % 4)), index=seed + i)[0]
            for i, n in enumerate(orders)]
    return capi.pack_batch(mats)


def options_for(orders):
    """Per-matrix options: fixed sweeps, the tolerance stop, the rate stop."""
    opts = []
    for i in range(len(orders)):
        k = i % 3
        if k == 0:
            opts.append(capi.jacobi_options(fixed_sweeps=3 + i % 5))
        elif k == 1:
            opts.append(capi.jacobi_options(off_tol=1e-7))
        else:
            opts.append(capi.jacobi_options(off_tol=1e-7, correction_rate=0.5))
    return opts


def run_both(runtime, orders, flat, options):
    _, _, _, total, vtotal = capi.batch_layout(orders)
    A = dev(flat)
    out = []
    for call in (runtime.jacobi_batched, runtime.jacobi_batched_noalloc):
        # Buffers filled with a pattern, so that an unwritten entry shows.
        V = torch.full((total,), 7.0, dtype=torch.float32, device="cuda")
        D = torch.full((vtotal,), 7.0, dtype=torch.float32, device="cuda")
        R = torch.full((vtotal,), 7.0, dtype=torch.float32, device="cuda")
        reps = call(orders, A, V, D, R, options)
        out.append((host(V), host(D), host(R), reps))
    return out


def same_reports(a, b):
    for ra, rb in zip(a, b):
        for name, _ in capi.JacobiReport._fields_:
            va, vb = getattr(ra, name), getattr(rb, name)
            if name == "off_history":
                assert list(va) == list(vb), name
            elif isinstance(va, float):
                assert np.float64(va).tobytes() == np.float64(vb).tobytes() or (
                    np.isnan(va) and np.isnan(vb)), name
            else:
                assert va == vb, name


@pytest.mark.parametrize("shape", ["mixed", "one_class"])
def test_noalloc_matches_allocating_entry_point(runtime, shape):
    if shape == "mixed":
        first = [1, 2, 5, 16, 17, 31, 32, 40, 48, 63, 64, 65, 80, 96, 100, 112, 127, 128,
                 129, 144, 150, 160, 176, 192, 200, 208, 224, 240, 255, 256] * 2
        second = [256, 3, 129, 64, 17, 200, 33]
    else:
        first = [64] * 40
        second = [64] * 9
    for label, orders, seed in (("first", first, 0), ("second", second, 100)):
        orders = list(orders)
        o, flat = batch(orders, seed)
        opts = options_for(o)
        say(f"noalloc {shape} {label}: {len(o)} matrices, orders {min(o)}..{max(o)}")
        (V1, D1, R1, r1), (V2, D2, R2, r2) = run_both(runtime, o, flat, opts)
        assert V1.tobytes() == V2.tobytes()
        assert D1.tobytes() == D2.tobytes()
        assert R1.tobytes() == R2.tobytes()
        same_reports(r1, r2)
        assert all(r.sweeps >= 1 for r in r2)
    # One options record for the whole batch.
    o, flat = batch([48, 200, 7], 300)
    (V1, D1, R1, r1), (V2, D2, R2, r2) = run_both(runtime, o, flat,
                                                  capi.jacobi_options(off_tol=1e-7))
    assert V1.tobytes() == V2.tobytes() and R1.tobytes() == R2.tobytes()
    same_reports(r1, r2)


def test_noalloc_rejects_what_its_scratch_cannot_hold(runtime):
    o, flat = batch([257], 0)
    _, _, _, total, vtotal = capi.batch_layout(o)
    A = dev(flat)
    V = torch.zeros(total, dtype=torch.float32, device="cuda")
    D = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    with pytest.raises(capi.MsqError, match="outside"):
        runtime.jacobi_batched_noalloc(o, A, V, D, D, capi.jacobi_options())
    too_many = [16] * (capi.NOALLOC_MAX_MATRICES + 1)
    with pytest.raises(capi.MsqError, match="count"):
        runtime.jacobi_batched_noalloc(too_many, A, V, D, D, capi.jacobi_options())
