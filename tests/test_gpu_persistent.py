"""The persistent Jacobi launches of orders 129..256 (one workgroup per CU
taking jobs from a device counter), scalar and blocked eigensolvers: more
jobs than workgroups, mixed orders,
every result equal bit for bit to the same matrix solved alone (a grid of
one), to the batch's graph replay, and close to LAPACK's eigenvalues."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, jacobi_method, say
from msq import capi
from msq import generators as G

pytestmark = pytest.mark.gpu


def jacobi(runtime, mats, opt):
    orders, flat = capi.pack_batch(mats)
    _, _, voffs, total, vtotal = capi.batch_layout(orders)
    V = torch.zeros(total, dtype=torch.float32, device="cuda")
    D = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    R = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    reps = runtime.jacobi_batched(orders, dev(flat), V, D, R, opt)
    Vs = capi.unpack_batch(orders, host(V))
    d = host(D)
    return Vs, [d[o:o + n] for o, n in zip(voffs, orders)], [r.sweeps for r in reps]


@pytest.fixture(params=["scalar", "blocked"])
def method(request, runtime):
    m = capi.JACOBI_SCALAR if request.param == "scalar" else capi.JACOBI_BLOCKED_MFMA
    with jacobi_method(runtime, m):
        yield runtime


@pytest.mark.parametrize("fixed", [0, 3])
def test_queue_with_more_jobs_than_workgroups(method, fixed):
    runtime = method
    rng = np.random.default_rng(41)
    orders = rng.integers(129, 257, size=600).tolist() + [16, 64, 100, 128]
    rng.shuffle(orders)
    mats = [G.single_matrix(9, int(n), 10.0 ** -rng.uniform(1, 4), index=i)[0]
            for i, n in enumerate(orders)]
    opt = capi.jacobi_options(off_tol=1e-7, max_sweeps=30, stagnation=0.5, floor_tol=1e-5,
                              fixed_sweeps=fixed)
    say(f"persistent queue: {len(orders)} matrices, fixed sweeps {fixed}")
    Vb, db, sb = jacobi(runtime, mats, opt)
    pick = [i for i, n in enumerate(orders) if n > 128][:12] + [orders.index(128)]
    for i in pick:
        Va, da, sa = jacobi(runtime, [mats[i]], opt)
        assert sa[0] == sb[i]
        assert np.array_equal(Va[0], Vb[i]) and np.array_equal(da[0], db[i]), (i, orders[i])
    for i in pick[:6] if not fixed else []:
        # converged solves (off(A) <= 1e-7 ||A||_F, or the FP32 floor at
        # 1e-5 ||A||_F): eigenvalues within that of LAPACK's
        lam = np.linalg.eigvalsh(mats[i])
        got = np.sort(db[i].astype(np.float64))
        assert np.max(np.abs(got - lam)) <= 1e-5 * np.linalg.norm(mats[i]), (i, orders[i])
        Vi = Vb[i].astype(np.float64)
        assert np.linalg.norm(Vi.T @ Vi - np.eye(orders[i])) <= 1e-4


def test_graph_replay_of_the_persistent_launch(method):
    runtime = method
    """The counter is cleared by a kernel inside the captured work, so every
    replay runs the whole queue again."""
    rng = np.random.default_rng(5)
    orders = rng.integers(129, 257, size=300).tolist()
    mats = [G.single_matrix(9, int(n), 1e-3, index=i)[0] for i, n in enumerate(orders)]
    o, flat = capi.pack_batch(mats)
    _, _, _, total, _ = capi.batch_layout(o)
    A = dev(flat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    from msq import solver as MS
    opts = MS.tier_options("bf16")
    runtime.set_graphs(True)
    try:
        outs = []
        for call in range(3):
            X.zero_()
            reps = runtime.solve_batched(o, A, X, opts)
            outs.append((host(X).copy(), reps.status.copy()))
    finally:
        runtime.set_graphs(False)
    for x, st in outs[1:]:
        assert np.array_equal(x, outs[0][0]) and np.array_equal(st, outs[0][1])
    assert (outs[0][1] == capi.CONVERGED).all()
