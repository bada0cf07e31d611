"""The blocked eigensolver of orders 17..256 (msq_runtime_set_jacobi_method
MSQ_JACOBI_BLOCKED_MFMA: 8-index blocks paired on the circle, each 16 x 16
pair subproblem rotated by scalar rotations, the rest of the round as FP32
MFMA products, V in registers), through the C ABI:

* sweep by sweep against the binary64 oracle of the same schedule
  (oracle.solver.jacobi_tile): diagonal, V and off(A) after 1 and 2 sweeps;
* converged solves against LAPACK: eigenvalues, orthogonality of V, V^T A V;
* a matrix's results do not depend on its batch (persistent queue), nor on
  graph capture and replay;
* whole solves with the blocked eigensolver meet 1e-6 on the batched uneven
  workload's orders 129..256 and on a sample of all its orders.
"""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, jacobi_method, say
from msq import capi
from msq import generators as G
from msq import solver as MS
from oracle import solver as S

pytestmark = pytest.mark.gpu


@pytest.fixture
def blocked(runtime):
    with jacobi_method(runtime, capi.JACOBI_BLOCKED_MFMA):
        yield runtime


def jacobi(runtime, mats, opt):
    orders, flat = capi.pack_batch(mats)
    _, _, voffs, total, vtotal = capi.batch_layout(orders)
    V = torch.zeros(total, dtype=torch.float32, device="cuda")
    D = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    R = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    reps = runtime.jacobi_batched(orders, dev(flat), V, D, R, opt)
    Vs = capi.unpack_batch(orders, host(V))
    d = host(D)
    return Vs, [d[o:o + n] for o, n in zip(voffs, orders)], reps


def graded(n, seed):
    """A well-separated, diagonally dominant SPD matrix: a permuted diagonal
    1..2 plus a symmetric perturbation of 1e-3.  Every rotation angle is far
    from the +-45 degree branch (a_qq - a_pp is never near 0), so a sweep in
    FP32 follows the binary64 oracle to rounding; on random dense matrices
    the first sweeps branch on near-ties of a_qq - a_pp, and runs differing by
    rounding part ways (the binary64 oracle itself does, for a 1e-7
    perturbation of A).  FP32 rounding still moves the rotations of the
    closest diagonal pairs (spacing 1 / n) a little: V follows to about
    1e-3 (a binary64 run on A plus 1e-6 noise differs by that much), the
    diagonal and off(A) to about 1e-7 and 1e-5 relative."""
    rng = np.random.default_rng(seed)
    E = rng.standard_normal((n, n)) * 1e-3
    return np.diag(rng.permutation(np.linspace(1.0, 2.0, n))) + (E + E.T) / 2


@pytest.mark.parametrize("sweeps", [1, 2])
def test_sweeps_follow_the_oracle(blocked, sweeps):
    orders = [17, 40, 100, 129, 144, 200, 256]
    mats = [graded(n, i) for i, n in enumerate(orders)]
    say(f"tile jacobi vs oracle: orders {orders}, {sweeps} sweeps")
    Vs, ds, reps = jacobi(blocked, mats, capi.jacobi_options(fixed_sweeps=sweeps))
    for A, V, d, rep in zip(mats, Vs, ds, reps):
        Vo, do, info = S.jacobi_tile(A, sweeps=sweeps, return_history=True)
        scale = np.linalg.norm(A)
        ed = np.max(np.abs(d.astype(np.float64) - do)) / scale
        ev = np.max(np.abs(V.astype(np.float64) - Vo))
        hist = np.array(rep.off_history[:sweeps], dtype=np.float64)
        eh = np.max(np.abs(hist - np.array(info["history"]) / scale) / hist)
        say(f"  n {A.shape[0]}: diag {ed:.2e}, V {ev:.2e}, off history {eh:.2e}")
        assert rep.sweeps == sweeps
        assert ed <= 1e-6 and ev <= 5e-3 and eh <= 1e-3


@pytest.mark.parametrize("sweeps", [1, 3])
def test_sweeps_are_similarity_transforms(blocked, sweeps):
    """Random dense matrices after a few sweeps: V orthogonal, diag(V^T A V)
    equal to the returned diagonal, off(V^T A V) to the reported off(A)."""
    orders = [20, 64, 129, 176, 256]
    mats = [G.single_matrix(9, n, 10.0 ** -(1 + i % 3), index=i)[0] for i, n in enumerate(orders)]
    Vs, ds, reps = jacobi(blocked, mats, capi.jacobi_options(fixed_sweeps=sweeps))
    for A, V, d, rep in zip(mats, Vs, ds, reps):
        n = A.shape[0]
        V64 = V.astype(np.float64)
        T = V64.T @ A @ V64
        scale = np.linalg.norm(A)
        eo = np.linalg.norm(V64.T @ V64 - np.eye(n))
        ed = np.max(np.abs(np.diag(T) - d)) / scale
        off = np.linalg.norm(T - np.diag(np.diag(T)))
        eoff = abs(off - rep.off_norm) / scale
        say(f"  n {n}, {sweeps} sweeps: orth {eo:.2e}, diag {ed:.2e}, off {eoff:.2e}")
        assert eo <= 1e-4 and ed <= 1e-5 and eoff <= 1e-5


def test_converged_solves_against_lapack(blocked):
    rng = np.random.default_rng(3)
    orders = [17, 33, 97, 129, 160, 191, 224, 255, 256]
    mats = [G.single_matrix(9, n, 10.0 ** -rng.uniform(1, 5), index=i)[0]
            for i, n in enumerate(orders)]
    opt = capi.jacobi_options(off_tol=1e-7, max_sweeps=30, stagnation=0.5, floor_tol=1e-5)
    Vs, ds, reps = jacobi(blocked, mats, opt)
    for A, V, d, rep in zip(mats, Vs, ds, reps):
        n = A.shape[0]
        lam = np.linalg.eigvalsh(A)
        scale = np.linalg.norm(A)
        V64 = V.astype(np.float64)
        el = np.max(np.abs(np.sort(d.astype(np.float64)) - lam)) / scale
        eo = np.linalg.norm(V64.T @ V64 - np.eye(n))
        T = V64.T @ A @ V64
        eoff = np.linalg.norm(T - np.diag(np.diag(T))) / scale
        say(f"  n {n}: sweeps {rep.sweeps}, off {rep.off_norm / rep.matrix_norm:.2e}, "
            f"eig {el:.2e}, orth {eo:.2e}, off(V^T A V) {eoff:.2e}")
        assert rep.invalid_spectrum == 0 and rep.sweeps < 30
        assert el <= 1e-5 and eo <= 1e-4 and eoff <= 2e-5


def test_results_do_not_depend_on_the_batch(blocked):
    rng = np.random.default_rng(41)
    orders = rng.integers(129, 257, size=400).tolist() + [16, 64, 100, 128]
    rng.shuffle(orders)
    mats = [G.single_matrix(9, int(n), 10.0 ** -rng.uniform(1, 4), index=i)[0]
            for i, n in enumerate(orders)]
    opt = capi.jacobi_options(off_tol=1e-7, max_sweeps=30, stagnation=0.5, floor_tol=1e-5)
    say(f"tile jacobi queue: {len(orders)} matrices")
    Vb, db, rb = jacobi(blocked, mats, opt)
    pick = [i for i, n in enumerate(orders) if n > 128][:10] + [orders.index(o) for o in (64, 100, 128)]
    for i in pick:
        Va, da, ra = jacobi(blocked, [mats[i]], opt)
        assert ra[0].sweeps == rb[i].sweeps
        assert np.array_equal(Va[0], Vb[i]) and np.array_equal(da[0], db[i]), (i, orders[i])


@pytest.mark.parametrize("low", [16, 128])
def test_solves_meet_the_target_and_replay_bit_for_bit(blocked, low):
    mats = [m for m in G.batched_uneven(count=4096) if m["order"] > low][:240]
    orders = [m["order"] for m in mats]
    o, aflat = capi.pack_batch([m["A"] for m in mats])
    lds, offs, _, total, _ = capi.batch_layout(o)
    A = dev(aflat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    blocked.set_graphs(True)
    try:
        outs = []
        for call in range(3):  # eager, captured, replayed
            X.zero_()
            reps = blocked.solve_batched(o, A, X, MS.tier_options("bf16"))
            outs.append((host(X).copy(), reps.status.copy(), reps.sweeps.copy()))
    finally:
        blocked.set_graphs(False)
    for x, st, sw in outs[1:]:
        assert np.array_equal(x, outs[0][0]) and np.array_equal(st, outs[0][1])
    Xf = outs[0][0]
    err = np.array([np.linalg.norm(Xf[off:off + ld * ld].reshape(ld, ld, order="F")[:n, :n]
                                   - m["X"]) / np.linalg.norm(m["X"])
                    for m, off, ld, n in zip(mats, offs, lds, orders)])
    say(f"tile solves: {len(mats)} matrices, mean sweeps {outs[0][2].mean():.2f}, "
        f"worst error {err.max():.3e}")
    assert (outs[0][1] == capi.CONVERGED).all()
    assert err.max() <= 1e-6
