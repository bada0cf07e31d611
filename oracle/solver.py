"""CPU oracle for the SPD square-root path (test infrastructure only).

This module restates, in numpy, the reference algorithm that the HIP library
implements.  It is the checker for the parity tests, the body of the CPU
baseline that ``bench.py`` times beside the GPU, and nothing else: the product
library never imports it.

Functions and the reference code they follow:

* ``jacobi_schedule``     -- makeJacobiSchedule, src/metal_solver.cpp:50-79
* ``jacobi_f32``          -- symmetricJacobi, src/metal_runtime.cpp:584-646 with
                             jacobi_rotations_f32 / jacobi_similarity_vectors_f32,
                             metal_mixed_sqrt.metal:1331-1411 (whole round applied
                             out of place, as the reference's kernels do)
* ``jacobi_tile``         -- the same rotations on blocks of 8 indices paired on the
                             circle (the HIP library's blocked eigensolver of orders
                             129..256): each 16 x 16 pair subproblem rotated by the
                             formula of jacobi_rotations_f32 (all its pairs in a
                             sweep's first round, its cross-block pairs after), the
                             rest of the round applied as Q^T A Q; binary64 (no
                             reference counterpart: the reference sweeps single
                             pairs only)
* ``frozen_factor``       -- prepareFrozenSpectralFactor,
                             src/metal_runtime.cpp:648-670, metal_mixed_sqrt.metal:1413-1434
* ``reconstruct_f32``     -- reconstructTiled, src/metal_runtime.cpp:429-459
* ``residual_fp64``       -- computeWorkingResidual, src/metal_runtime.cpp:789-834
                             (float-float there, binary64 here)
* ``correction_f32``      -- refineUpdateAndMeasure FP32 branch,
                             src/metal_runtime.cpp:867-889, plus
                             working_add_symmetric_correction (:1955-1970 of the .metal)
* ``solve_gpu_resident``  -- solveGpuResidentSpd, src/metal_solver.cpp:92-336
* ``reference_root``      -- symmetricReferenceRoot, src/accelerate_backend.cpp:594-607
                             (LAPACK dsyevd through numpy.linalg.eigh)
* ``relative_error`` etc. -- src/numerics.cpp:213-264
* ``accelerate_mixed``    -- the CPU plain path, 1_baseline/spd.cpp:348-430
                             (LAPACK ssyevd/sgemm/dgemm through scipy)
* ``controlled_matrix``   -- makeSpdMatrix, src/numerics.cpp:141-195

Pinning (what keeps this restatement honest):
* schedule: every round is a perfect matching and every pair appears exactly
  once per sweep (closed-form round-robin property), checked in tests;
* Jacobi: eigenvalues against LAPACK, orthogonality of V, off-norm decrease;
* jacobi_tile: at order 16 (one block pair) a sweep is one scalar sweep of the
  circle schedule, checked against jacobi_f32; at larger orders eigenvalues
  against LAPACK, orthogonality, V^T A V against the rotated A;
* products/residual/correction: closed forms against float64 numpy, and the
  reference's own known answer (identity basis, rhs 2, H = 0.5 gives 1
  exactly, src/metal_checks.cpp:537-601);
* end to end: forward error against the LAPACK root and against the
  construction-known root Q diag(sqrt(lambda)) Q^T.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
F64 = np.float64


# ----------------------------------------------------------------------------
# Round-robin schedule
# ----------------------------------------------------------------------------
def jacobi_schedule(order: int):
    """Circle-method disjoint pairs (src/metal_solver.cpp:50-79).

    Returns (padded_order, partners, pair_indices, pair_members), each of the
    three tables shaped (rounds, padded_order) as in the reference.
    """
    padded = order + order % 2
    rounds = padded - 1
    participants = list(range(padded))
    partners = np.zeros((rounds, padded), dtype=np.uint32)
    pair_indices = np.zeros((rounds, padded), dtype=np.uint32)
    pair_members = np.zeros((rounds, padded), dtype=np.uint32)
    for r in range(rounds):
        for pair in range(padded // 2):
            first = participants[pair]
            second = participants[padded - 1 - pair]
            pair_members[r, 2 * pair] = min(first, second)
            pair_members[r, 2 * pair + 1] = max(first, second)
            partners[r, first] = second
            partners[r, second] = first
            pair_indices[r, first] = pair
            pair_indices[r, second] = pair
        # std::rotate(begin + 1, end - 1, end): last element moves to slot 1.
        participants = [participants[0], participants[-1]] + participants[1:-1]
    return padded, partners, pair_indices, pair_members


# ----------------------------------------------------------------------------
# FP32 parallel cyclic Jacobi
# ----------------------------------------------------------------------------
def _rotations(A, p, q, n, angle_dtype):
    """Rotation (c, s) per pair, jacobi_rotations_f32 (.metal:1331-1361).

    ``angle_dtype`` float32 restates the reference exactly; float64 computes
    tau, t, c, s in binary64 and rounds c and s to float32 once (the HIP
    library's choice).
    """
    valid = (p < n) & (q < n)
    pp = np.where(valid, p, 0)
    qq = np.where(valid, q, 0)
    app = A[pp, pp].astype(angle_dtype)
    aqq = A[qq, qq].astype(angle_dtype)
    apq = (angle_dtype(0.5) * (A[pp, qq].astype(angle_dtype) +
                               A[qq, pp].astype(angle_dtype)))
    c = np.ones_like(app)
    s = np.zeros_like(app)
    nz = valid & (apq != 0)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        tau = (aqq - app) / (angle_dtype(2.0) * apq)
        t = np.copysign(angle_dtype(1.0) /
                        (np.abs(tau) + np.sqrt(angle_dtype(1.0) + tau * tau)), tau)
        cc = angle_dtype(1.0) / np.sqrt(angle_dtype(1.0) + t * t)
        ss = t * cc
    c = np.where(nz, cc, c).astype(F32)
    s = np.where(nz, ss, s).astype(F32)
    return c, s


def jacobi_f32(A, sweeps=8, off_tolerance=None, max_sweeps=None,
               angle_dtype=F64, return_history=False, pad_to=None):
    """Parallel cyclic Jacobi on a float32 copy of A.

    sweeps          -- fixed sweep count (the reference runs 8,
                       src/metal_solver.cpp:210-213) when off_tolerance is None.
    off_tolerance   -- stop after the first sweep whose FP64 off-diagonal norm
                       is <= off_tolerance * ||A||_F (new: the reference has no
                       convergence test); at most max_sweeps sweeps.

    pad_to          -- run the schedule of this (even) order, with zero rows
                       and columns beyond n (the HIP library pads orders to a
                       multiple of 16; for orders that are already multiples
                       of 16 the schedule is the reference's).

    Returns (V, root_eigenvalues, info) with V float32 column-major-equivalent
    (V[:, j] is the j-th eigenvector) and root_eigenvalues = sqrt(diag) as in
    jacobi_root_eigenvalues_f32 (.metal:1436-1444), unsorted.
    """
    A = np.array(A, dtype=F32, copy=True)
    n_real = A.shape[0]
    if pad_to is not None and pad_to > n_real:
        Ap = np.zeros((pad_to, pad_to), dtype=F32)
        Ap[:n_real, :n_real] = A
        V, roots, info = jacobi_f32(Ap, sweeps=sweeps, off_tolerance=off_tolerance,
                                    max_sweeps=max_sweeps, angle_dtype=angle_dtype,
                                    return_history=return_history)
        info["diagonal"] = info["diagonal"][:n_real]
        return V[:n_real, :n_real], roots[:n_real], info
    n = n_real
    V = np.eye(n, dtype=F32)
    padded, partners, pair_indices, pair_members = jacobi_schedule(n)
    rows = np.arange(n)
    norm_a = float(np.sqrt(np.sum(np.asarray(A, F64) ** 2)))
    if off_tolerance is None:
        total = sweeps
    else:
        total = max_sweeps if max_sweeps is not None else 30
    history = []
    done = 0
    for sweep in range(total):
        for r in range(padded - 1):
            p = pair_members[r, 0::2].astype(np.int64)
            q = pair_members[r, 1::2].astype(np.int64)
            c, s = _rotations(A, p, q, n, angle_dtype)
            rot = np.zeros((padded, 2), dtype=F32)
            rot[0::2, 0] = c
            rot[0::2, 1] = -s
            rot[1::2, 0] = c
            rot[1::2, 1] = s
            part = partners[r, :n].astype(np.int64)
            pidx = pair_indices[r, :n].astype(np.int64)
            slot = 2 * pidx + (part < rows)
            rr = rot[slot]                      # rotation seen by each index
            safe = np.where(part < n, part, rows)
            # A' = J^T A J, element formula of .metal:1394-1410
            rx = rr[:, 0][:, None]
            ry = rr[:, 1][:, None]
            cx = rr[:, 0][None, :]
            cy = rr[:, 1][None, :]
            A_new = (rx * cx * A + ry * cx * A[safe, :] +
                     rx * cy * A[:, safe] + ry * cy * A[np.ix_(safe, safe)])
            # V' = V J, z == 1 slice of .metal:1375-1384
            V_new = np.where((part < n)[None, :],
                             rr[:, 0][None, :] * V + rr[:, 1][None, :] * V[:, safe],
                             V)
            A = A_new.astype(F32)
            V = V_new.astype(F32)
        done += 1
        Ad = np.asarray(A, F64)
        off = float(np.sqrt(max(np.sum(Ad * Ad) - np.sum(np.diag(Ad) ** 2), 0.0)))
        history.append(off)
        if off_tolerance is not None and off <= off_tolerance * norm_a:
            break
    with np.errstate(invalid="ignore"):
        roots = np.sqrt(np.diag(A)).astype(F32)
    info = {"sweeps": done, "off_norm": history[-1] if history else 0.0,
            "matrix_norm": norm_a, "diagonal": np.diag(A).copy()}
    if return_history:
        info["history"] = history
    return V, roots, info


def _circle(t, x, np_):
    """Participant at slot x of round t (closed form of makeJacobiSchedule)."""
    return 0 if x == 0 else 1 + ((x - 1 - t) % (np_ - 1))


def _angles64(app, aqq, apq):
    """(c, s) of the pair p < q in binary64: t = sign(d e) |e| / (|d| +
    sqrt(d^2 + e^2)), d = a_qq - a_pp, e = 2 a_pq, c = 1 / sqrt(1 + t^2),
    s = t c; (1, 0) where a_pq = 0 (.metal:1331-1361 in binary64)."""
    d = aqq - app
    e = 2.0 * apq
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.copysign(np.abs(e) / (np.abs(d) + np.sqrt(d * d + e * e)), d * e)
        c = 1.0 / np.sqrt(1.0 + t * t)
    nz = apq != 0
    return np.where(nz, c, 1.0), np.where(nz, t * c, 0.0)


def jacobi_tile(A, sweeps=1, off_tolerance=None, max_sweeps=30, return_history=False):
    """Blocked cyclic Jacobi in binary64, the schedule of the HIP library's
    blocked eigensolver.

    The order is padded to N, a multiple of 16 (zero rows and columns), and
    cut into NB = N / 8 blocks of 8 indices.  Round t of a sweep (NB - 1
    rounds) pairs block P_t(k) with block P_t(NB - 1 - k) on the circle; the
    pair's 16 indices (local 0..7 the first block, 8..15 the second) form the
    subproblem S = A[pair, pair].  In round 0 it is swept by the circle
    schedule over its 16 indices (15 rounds of 8 disjoint rotations), in the
    other rounds over the 64 pairs across its two blocks only (8 rounds, local
    u with 8 + (7 - u - r) mod 8 in round r): every pair of indices is rotated
    exactly once per sweep, as in a scalar sweep.  Each rotation is A' = J^T A
    J with J_pp = J_qq = c, J_pq = s, J_qp = -s, accumulated as Q <- Q J.  Then
    A <- J^T A J and V <- V J for the block-diagonal J of the round's Q.

    Returns (V, diagonal, info) like jacobi_f32, in binary64."""
    A0 = np.asarray(A, F64)
    n = A0.shape[0]
    N = (n + 15) // 16 * 16
    NB, M = N // 8, N // 16
    B = np.zeros((N, N))
    B[:n, :n] = A0
    V = np.eye(N)
    norm_a = float(np.linalg.norm(A0))
    total = sweeps if off_tolerance is None else max_sweeps
    history = []
    done = 0
    for _ in range(total):
        for t in range(NB - 1):
            idx = np.array([np.r_[8 * _circle(t, k, NB):8 * _circle(t, k, NB) + 8,
                                  8 * _circle(t, NB - 1 - k, NB):
                                  8 * _circle(t, NB - 1 - k, NB) + 8] for k in range(M)])
            S = B[idx[:, :, None], idx[:, None, :]].copy()
            Q = np.tile(np.eye(16), (M, 1, 1))
            if t == 0:  # every pair of the 16 indices: the circle over 16
                inner = [[(_circle(tau, u, 16), _circle(tau, 15 - u, 16)) for u in range(8)]
                         for tau in range(15)]
            else:  # the pairs across the two blocks only
                inner = [[(u, 8 + (7 - u - r) % 8) for u in range(8)] for r in range(8)]
            for pairs in inner:
                for a, b in pairs:
                    p, q = min(a, b), max(a, b)
                    c, s = _angles64(S[:, p, p], S[:, q, q], S[:, p, q])
                    c, s = c[:, None], s[:, None]
                    sp, sq = S[:, p, :].copy(), S[:, q, :].copy()
                    S[:, p, :], S[:, q, :] = c * sp - s * sq, s * sp + c * sq
                    sp, sq = S[:, :, p].copy(), S[:, :, q].copy()
                    S[:, :, p], S[:, :, q] = c * sp - s * sq, s * sp + c * sq
                    qp, qq = Q[:, :, p].copy(), Q[:, :, q].copy()
                    Q[:, :, p], Q[:, :, q] = c * qp - s * qq, s * qp + c * qq
            J = np.zeros((N, N))
            for k in range(M):
                J[np.ix_(idx[k], idx[k])] = Q[k]
            B = J.T @ B @ J
            V = V @ J
        done += 1
        off = float(np.linalg.norm(B - np.diag(np.diag(B))))
        history.append(off)
        if off_tolerance is not None and off <= off_tolerance * norm_a:
            break
    info = {"sweeps": done, "off_norm": history[-1] if history else 0.0,
            "matrix_norm": norm_a, "diagonal": np.diag(B)[:n].copy(), "rotated": B[:n, :n]}
    if return_history:
        info["history"] = history
    return V[:n, :n], np.diag(B)[:n].copy(), info


# ----------------------------------------------------------------------------
# Frozen factor, products, residual, correction
# ----------------------------------------------------------------------------
def frozen_factor(V, roots):
    """V^T and H_ij = 1 / (s_i + s_j) (.metal:1413-1434)."""
    roots = np.asarray(roots, F32)
    H = (F32(1.0) / (roots[:, None] + roots[None, :])).astype(F32)
    return np.ascontiguousarray(np.asarray(V, F32).T), H


def reconstruct_f32(V, roots):
    """X0 = (V diag(s)) V^T with float32 operands (src/metal_runtime.cpp:429-459).

    Accumulated in float64 and rounded once: a tighter product than any FP32
    accumulation order, so GPU results are compared against it with an FP32
    accumulation bound.
    """
    V = np.asarray(V, F64)
    W = (V * np.asarray(roots, F64)[None, :]).astype(F32).astype(F64)
    return (W @ V.T).astype(F32)


def symmetrize(X):
    X = np.asarray(X, F64)
    return 0.5 * (X + X.T)


def residual_fp64(A, X):
    """R = A - X X and ||R||_F / ||A||_F in binary64 (src/metal_runtime.cpp:789-834)."""
    A = np.asarray(A, F64)
    X = np.asarray(X, F64)
    R = A - X @ X
    return R, float(np.sqrt(np.sum(R * R) / np.sum(A * A)))


def correction_f32(V, H, R):
    """One frozen-factor correction E (src/metal_runtime.cpp:867-889).

    W = V^T R; Y = (W V) o H; W = V Y; E = W V^T, operands float32, every
    product accumulated in float64 and rounded to float32 (see reconstruct_f32).
    """
    V64 = np.asarray(V, F32).astype(F64)
    R32 = np.asarray(R, F64).astype(F32).astype(F64)
    W = (V64.T @ R32).astype(F32).astype(F64)
    Y = ((W @ V64).astype(F32) * np.asarray(H, F32)).astype(F32).astype(F64)
    W = (V64 @ Y).astype(F32).astype(F64)
    E = (W @ V64.T).astype(F32)
    return E


def apply_correction(X, E):
    """X += (E + E^T) / 2 (working_add_symmetric_correction, .metal:1955-1970)."""
    E = np.asarray(E, F64)
    return np.asarray(X, F64) + 0.5 * (E + E.T)


def solve_gpu_resident(A, sweeps=8, off_tolerance=None, max_sweeps=None,
                       max_iterations=12, tolerance=1e-12, basis_quantizer=None,
                       angle_dtype=F64):
    """solveGpuResidentSpd (src/metal_solver.cpp:92-336), FP32 stable basis.

    basis_quantizer -- optional callable V -> decoded low-precision V; when given,
    X0 is reconstructed from it (reconstructLowFactorTiled, :286-288) and the
    corrections switch back to the FP32 basis (:304-310), as the reference does.
    """
    A = np.asarray(A, F64)
    V, roots, info = jacobi_f32(A, sweeps=sweeps, off_tolerance=off_tolerance,
                                max_sweeps=max_sweeps, angle_dtype=angle_dtype)
    Vx = V if basis_quantizer is None else basis_quantizer(V)
    X = symmetrize(reconstruct_f32(Vx, roots))
    X0 = X.copy()
    _, H = frozen_factor(V, roots)
    R, res = residual_fp64(A, X)
    residuals = [res]
    for _ in range(max_iterations):
        if not res > tolerance:
            break
        E = correction_f32(V, H, R)
        X = apply_correction(X, E)
        R, res = residual_fp64(A, X)
        residuals.append(res)
        if not np.isfinite(res):
            break
    return X, {"residuals": residuals, "initial_root": X0, "V": V,
               "roots": roots, "jacobi": info}


# ----------------------------------------------------------------------------
# FP64 oracle and metrics
# ----------------------------------------------------------------------------
def reference_root(A):
    """symmetricReferenceRoot (src/accelerate_backend.cpp:594-607).

    LAPACK dsyevd (numpy.linalg.eigh), V diag(sqrt(lambda)) V^T, symmetrised.
    """
    A = np.asarray(A, F64)
    w, Q = np.linalg.eigh(A)
    if not np.all(w > 0):
        raise ValueError("The FP64 reference has a nonpositive eigenvalue.")
    X = (Q * np.sqrt(w)[None, :]) @ Q.T
    return 0.5 * (X + X.T)


def relative_error(X, Xref):
    """relativeError (src/numerics.cpp:230-240)."""
    X = np.asarray(X, F64)
    Xref = np.asarray(Xref, F64)
    return float(np.sqrt(np.sum((X - Xref) ** 2) / np.sum(Xref ** 2)))


def relative_residual(A, X):
    """relativeResidual (src/numerics.cpp:213-228)."""
    return residual_fp64(A, X)[1]


def relative_symmetry_error(X):
    """relativeSymmetryError (src/numerics.cpp:252-264)."""
    X = np.asarray(X, F64)
    return float(np.sqrt(np.sum((X - X.T) ** 2) / np.sum(X * X)))


def orthogonality_error(V):
    """||V^T V - I||_F / sqrt(n) (src/metal_runtime.cpp:742-776)."""
    V = np.asarray(V, F64)
    n = V.shape[0]
    G = V.T @ V - np.eye(n)
    return float(np.sqrt(np.sum(G * G) / n))


# ----------------------------------------------------------------------------
# Reference matrix family and the CPU plain path
# ----------------------------------------------------------------------------
def controlled_matrix(order, condition):
    """makeSpdMatrix (src/numerics.cpp:141-195): sin/cos Gram-Schmidt basis."""
    n = order
    Q = np.zeros((n, n))
    r = np.arange(n, dtype=F64)
    for c in range(n):
        v = np.sin(0.173 * (r + 1) * (c + 2)) + np.cos(0.117 * (r + 3) * (c + 1))
        for prior in range(c):
            v = v - np.dot(Q[:, prior], v) * Q[:, prior]
        v = v / np.sqrt(np.dot(v, v))
        Q[:, c] = v
    lam = np.exp(np.log(condition) * np.arange(n) / max(1, n - 1))
    return (Q * lam[None, :]) @ Q.T


def accelerate_mixed(A, Xref=None, max_corrections=12, tolerance=1e-12):
    """CPU plain path, 1_baseline/spd.cpp:348-430.

    FP32 ssyevd, sgemm reconstruction, FP64 dgemm residual, FP32 corrections
    with four sgemm and the diagonal Sylvester solve in the eigenbasis.
    """
    from scipy.linalg import blas, lapack

    A = np.asarray(A, F64)
    n = A.shape[0]
    w, Q, info = lapack.ssyevd(np.asfortranarray(A.astype(F32)), compute_v=1,
                               lower=0)
    if info != 0 or not np.all(w > 0):
        raise ValueError("ssyevd failed or returned a nonpositive eigenvalue")
    lam = np.sqrt(w).astype(F32)
    Q = np.asfortranarray(Q.astype(F32))
    scaled = np.asfortranarray(Q * lam[None, :])
    X = blas.sgemm(1.0, scaled, Q, trans_b=1).astype(F64)
    X = 0.5 * (X + X.T)
    P = blas.dgemm(1.0, np.asfortranarray(X), np.asfortranarray(X))
    R = A - P
    res = float(np.sqrt(np.sum(R * R) / np.sum(A * A)))
    steps = [(res, relative_error(X, Xref) if Xref is not None else np.nan)]
    denom = lam[:, None] + lam[None, :]
    for _ in range(max_corrections):
        if not res > tolerance:
            break
        R32 = np.asfortranarray(R.astype(F32))
        W = blas.sgemm(1.0, Q, R32, trans_a=1)
        C = blas.sgemm(1.0, W, Q)
        C = (F32(0.5) * (C + C.T)) / denom
        W = blas.sgemm(1.0, Q, np.asfortranarray(C.astype(F32)))
        E = blas.sgemm(1.0, W, Q, trans_b=1)
        X = X + E.astype(F64)
        X = 0.5 * (X + X.T)
        P = blas.dgemm(1.0, np.asfortranarray(X), np.asfortranarray(X))
        R = A - P
        res = float(np.sqrt(np.sum(R * R) / np.sum(A * A)))
        steps.append((res, relative_error(X, Xref) if Xref is not None else np.nan))
    return X, steps
