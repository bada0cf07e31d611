"""Solver configurations of the SPD square-root path and the tiered solve.

The reference's configuration is MetalSolverConfiguration / MetalProblem
(src/metal_solver.hpp:12-31): basis and spectrum precisions, a fixed 8-sweep
Jacobi (src/metal_solver.cpp:210-213), corrections in FP32 on the stable basis
until the residual tolerance (default 1e-12, src/metal_backend.cpp:105-113).
Here a configuration is an ``msq_solve_options`` record per matrix:

* ``plain_fp32_options`` -- the plain FP32 GPU path: FP32 Jacobi to
  off(A) <= 1e-7 ||A||_F (or until it stalls at the FP32 floor), FP32
  reconstruction and corrections, binary64 residual; corrections stop when the
  forward-error estimate ||R||_F / (2 s_min ||X||_F) reaches the target.
* ``tier_options(name)`` -- a low-bit tier: reconstruction and corrections in
  BF16 / scaled FP8 / MXFP4 products, Jacobi stopped as soon as the frozen
  factor's corrections are predicted to contract by ``rate`` per step
  (||G||_F / min s, G_ij = F_ij / (s_i + s_j) with F the off-diagonal part of
  the rotated matrix), and more corrections allowed.
* ``solve_with_fallback`` -- runs tiered options and re-solves every matrix
  whose binary64 residual check missed the target on the plain FP32 path.
"""
from __future__ import annotations

import numpy as np

from . import capi

# Forward-error target of the path (relative Frobenius error vs FP64).
TARGET = 1e-6
# The estimate ||R|| / (2 s_min ||X||) bounds the error to first order; half
# the target leaves room for its second-order term.
ESTIMATE_TARGET = 0.5 * TARGET
# A matrix whose smallest Jacobi eigenvalue is below RESOLUTION times its
# largest is beyond what the FP32 Jacobi resolves: its estimate is not
# trusted (status "unresolved"), and the fallback solve runs the second
# Jacobi pass on V^T A V for it.  Measured on MI355X over 1000 matrices of
# the n in [32, 512] mixture (cond up to ~1e8): every matrix whose plain
# solve missed 1e-6 had a Jacobi eigenvalue ratio below 3e-7; 5e-7 keeps a
# margin and flags about 4.5% of the mixture.
RESOLUTION = 5e-7

TIER_PRECISIONS = {
    "fp32": (capi.FP32, capi.FP32),
    "bf16": (capi.BF16, capi.BF16),
    "fp8": (capi.FP8_E4M3FN_SCALED, capi.BF16),
    "mxfp4": (capi.MXFP4_E2M1, capi.BF16),
}


def plain_fp32_options(max_corrections=8, min_corrections=0, refine=False):
    """refine: the second Jacobi pass for unresolved spectra (the fallback)."""
    return capi.solve_options(
        capi.FP32, capi.FP32,
        capi.jacobi_options(off_tol=1e-7, max_sweeps=30, stagnation=0.5, floor_tol=1e-5),
        min_corrections=min_corrections, max_corrections=max_corrections,
        residual_tolerance=0.0, forward_tolerance=ESTIMATE_TARGET, resolution=RESOLUTION,
        refine_unresolved=1 if refine else 0)


def fallback_options():
    """The plain path with the second Jacobi pass, and room for the extra
    corrections of very ill-conditioned matrices."""
    return plain_fp32_options(max_corrections=30, refine=True)


def tier_options(name, rate=0.5, max_corrections=12, min_corrections=0):
    """A low-bit tier.  rate: Jacobi stops once one correction is predicted to
    shrink the root's error by this factor (first-order bound)."""
    basis, corr = TIER_PRECISIONS[name]
    return capi.solve_options(
        basis, corr,
        capi.jacobi_options(off_tol=1e-7, max_sweeps=30, stagnation=0.5, floor_tol=1e-5,
                            correction_rate=rate),
        min_corrections=min_corrections, max_corrections=max_corrections,
        residual_tolerance=0.0, forward_tolerance=ESTIMATE_TARGET, resolution=RESOLUTION)


def solve_with_fallback(runtime, orders, matrices, roots, options, packed_host=None):
    """Solve with per-matrix options; matrices that did not converge (an
    unresolved spectrum included) are solved again on the plain FP32 path with
    the second Jacobi pass.  Returns (reports, fallback_indices).

    matrices / roots: device tensors in batch layout.  The retry runs on the
    failed matrices only (gathered into a sub-batch)."""
    import torch

    reports = runtime.solve_batched(orders, matrices, roots, options)
    failed = np.flatnonzero(~reports.converged).tolist()
    if not failed:
        return reports, failed
    lds, offs, _, _, _ = capi.batch_layout(orders)
    sub_orders = [orders[i] for i in failed]
    _, soffs, _, stotal, _ = capi.batch_layout(sub_orders)
    sub_a = torch.zeros(stotal, dtype=torch.float64, device=matrices.device)
    sub_x = torch.zeros(stotal, dtype=torch.float64, device=matrices.device)
    for k, i in enumerate(failed):
        size = int(lds[i]) ** 2
        sub_a[soffs[k]:soffs[k] + size] = matrices[offs[i]:offs[i] + size]
    sub_reports = runtime.solve_batched(sub_orders, sub_a, sub_x, fallback_options())
    for k, i in enumerate(failed):
        size = int(lds[i]) ** 2
        roots[offs[i]:offs[i] + size] = sub_x[soffs[k]:soffs[k] + size]
        reports.replace(i, sub_reports, k)
    return reports, failed
