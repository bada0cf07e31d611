"""Seeded synthetic SPD inputs for the SPD square-root path.

A = Q diag(lambda) Q^T with Q Haar-random orthogonal (QR of a seeded N(0,1)
matrix, columns sign-fixed so that diag(R) > 0).  The construction-known root
Q diag(sqrt(lambda)) Q^T is returned beside A.

Random numbers come from xorshift64* uniforms and Box-Muller normals, the
generator of the reference (1_baseline/spd.cpp:137-168,
src/accelerate_backend.cpp:626-664).  To make generation vectorisable, every
column of the Gaussian matrix draws from its own xorshift64* stream; all
streams of one batch step in lock-step.

Seeds: matrix ``index`` of workload ``config`` uses
    seed = 0x9e3779b97f4a7c15 ^ (config << 32) ^ index,
its column j uses  seed ^ ((j + 1) * 0xbf58476d1ce4e5b9)  (mod 2^64), and its
scalar parameters (order, condition, spectrum) come from the stream ``seed``
itself.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_MIX = np.uint64(0xBF58476D1CE4E5B9)
_MULT = np.uint64(2685821657736338717)


def matrix_seed(config: int, index: int) -> int:
    return (0x9E3779B97F4A7C15 ^ (int(config) << 32) ^ int(index)) & (2**64 - 1)


class Streams:
    """A vector of independent xorshift64* streams stepped in lock-step."""

    def __init__(self, seeds):
        s = np.asarray(seeds, dtype=np.uint64).copy()
        s[s == 0] = _GOLDEN
        self.state = s
        self.spare = None

    def uniform(self):
        with np.errstate(over="ignore"):
            v = self.state
            v = v ^ (v >> np.uint64(12))
            v = v ^ (v << np.uint64(25))
            v = v ^ (v >> np.uint64(27))
            self.state = v
            bits = (v * _MULT) >> np.uint64(11)
        return bits.astype(np.float64) * 2.0**-53

    def normal(self):
        if self.spare is not None:
            out, self.spare = self.spare, None
            return out
        u = np.maximum(self.uniform(), np.finfo(np.float64).tiny)
        angle = 2.0 * np.pi * self.uniform()
        radius = np.sqrt(-2.0 * np.log(u))
        self.spare = radius * np.sin(angle)
        return radius * np.cos(angle)


def haar_orthogonal(seeds, order):
    """Haar-random orthogonal matrices, one per seed, shape (len(seeds), n, n)."""
    seeds = np.asarray(seeds, dtype=np.uint64)
    batch = seeds.shape[0]
    cols = np.arange(1, order + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        col_seeds = seeds[:, None] ^ (cols[None, :] * _MIX)
    st = Streams(col_seeds.reshape(-1))
    G = np.empty((order, batch * order))
    for row in range(order):
        G[row] = st.normal()
    G = G.reshape(order, batch, order).transpose(1, 0, 2)  # (batch, row, col)
    Q, R = np.linalg.qr(G)
    d = np.sign(np.diagonal(R, axis1=1, axis2=2))
    d[d == 0] = 1.0
    return Q * d[:, None, :]


def log_uniform_spectrum(u, order, condition):
    """lambda = cond^(u - 1) for u uniform, extremes pinned to 1/cond and 1."""
    lam = condition ** (u[:order] - 1.0)
    lam[0] = 1.0 / condition
    if order > 1:
        lam[1] = 1.0
    return lam


def clustered_spectrum(u, order, condition, clusters=4, spread=1e-3):
    """4 cluster centres log-uniform on [1/cond, 1]; each eigenvalue sits at a
    random centre times (1 + spread * (2u - 1)), clamped to [1/cond, 1];
    extremes pinned."""
    centres = condition ** (u[:clusters] - 1.0)
    centres[0] = 1.0 / condition
    centres[-1] = 1.0
    pick = np.minimum((u[clusters:clusters + order] * clusters).astype(int),
                      clusters - 1)
    jitter = 1.0 + spread * (2.0 * u[clusters + order:clusters + 2 * order] - 1.0)
    lam = np.clip(centres[pick] * jitter, 1.0 / condition, 1.0)
    lam[0] = 1.0 / condition
    if order > 1:
        lam[1] = 1.0
    return lam


def power_law_spectrum(order, alpha):
    """lambda_k = k^-alpha, k = 1..n."""
    return np.arange(1, order + 1, dtype=np.float64) ** (-alpha)


def assemble(Q, lam):
    """A = Q diag(lam) Q^T and X* = Q diag(sqrt(lam)) Q^T, both symmetrised."""
    A = (Q * lam[..., None, :]) @ np.swapaxes(Q, -1, -2)
    X = (Q * np.sqrt(lam)[..., None, :]) @ np.swapaxes(Q, -1, -2)
    A = 0.5 * (A + np.swapaxes(A, -1, -2))
    X = 0.5 * (X + np.swapaxes(X, -1, -2))
    return A, X


def _param_uniforms(seed, count):
    st = Streams([seed])
    return np.array([st.uniform()[0] for _ in range(count)])


def single_matrix(config, order, lam_min, index=0):
    """One matrix with log-uniform spectrum on [lam_min, 1] (configs 0-2)."""
    seed = matrix_seed(config, index)
    Q = haar_orthogonal([seed], order)[0]
    u = _param_uniforms(seed ^ 0x5DEECE66D, order)
    lam = log_uniform_spectrum(u, order, 1.0 / lam_min)
    A, X = assemble(Q, lam)
    return A, X, lam


def batched_uneven(count=4096, config=3, order_step=16, order_classes=16,
                   log_cond=(1.0, 5.0)):
    """The batched uneven workload.

    Matrix i: n = order_step * (1 + floor(order_classes * u0)),
    cond = 10^(log_cond[0] + (log_cond[1]-log_cond[0]) u1); even i log-uniform
    spectrum, odd i four tight clusters (relative spread 1e-3).
    Returns a list of dicts with keys order, condition, kind, A, X, lam, seed.
    """
    metas = []
    for i in range(count):
        seed = matrix_seed(config, i)
        u = _param_uniforms(seed, 2)
        n = order_step * (1 + min(int(order_classes * u[0]), order_classes - 1))
        cond = 10.0 ** (log_cond[0] + (log_cond[1] - log_cond[0]) * u[1])
        metas.append((i, seed, n, cond))
    out = [None] * count
    for n in sorted({m[2] for m in metas}):
        group = [m for m in metas if m[2] == n]
        Q = haar_orthogonal([m[1] for m in group], n)
        for k, (i, seed, _, cond) in enumerate(group):
            u = _param_uniforms(seed ^ 0x5DEECE66D, 2 * n + 4)
            if i % 2 == 0:
                lam = log_uniform_spectrum(u, n, cond)
                kind = "log_uniform"
            else:
                lam = clustered_spectrum(u, n, cond)
                kind = "clustered"
            A, X = assemble(Q[k], lam)
            out[i] = {"index": i, "order": n, "condition": cond, "kind": kind,
                      "A": A, "X": X, "lam": lam, "seed": seed}
    return out


def mixture(count=20000, config=4, order_range=(32, 512), max_log_cond=6.0,
            alpha_range=(0.5, 3.0)):
    """The learned-tier workload: n uniform in order_range, spectra one third
    each log-uniform (cond up to 10^max_log_cond), power law k^-alpha and four
    tight clusters.  Yields dicts as batched_uneven does, generated lazily in
    order of index (large batches do not have to live in memory at once)."""
    lo, hi = order_range
    for i in range(count):
        seed = matrix_seed(config, i)
        u = _param_uniforms(seed, 3)
        n = lo + min(int((hi - lo + 1) * u[0]), hi - lo)
        Q = haar_orthogonal([seed], n)[0]
        w = _param_uniforms(seed ^ 0x5DEECE66D, 2 * n + 4)
        kind_id = min(int(3 * u[1]), 2)
        if kind_id == 0:
            cond = 10.0 ** (max_log_cond * u[2])
            lam = log_uniform_spectrum(w, n, cond)
            kind = "log_uniform"
        elif kind_id == 1:
            alpha = alpha_range[0] + (alpha_range[1] - alpha_range[0]) * u[2]
            lam = power_law_spectrum(n, alpha)
            cond = float(lam.max() / lam.min())
            kind = "power_law"
        else:
            cond = 10.0 ** (max_log_cond * u[2])
            lam = clustered_spectrum(w, n, cond)
            kind = "clustered"
        A, X = assemble(Q, lam)
        yield {"index": i, "order": n, "condition": cond, "kind": kind,
               "A": A, "X": X, "lam": lam, "seed": seed}


# ----------------------------------------------------------- paper test sets
PAPER_ORDERS = {"V": 4096, "VI": 4096, "VII": 1024}
_PAPER_ORDINALS = {"V": 5, "VI": 6, "VII": 7}


class _ScalarStream:
    """One xorshift64* stream drawn value by value (uniform on [0, 1) with 53
    bits; Box-Muller normals in pairs, cos first, the sin value kept)."""

    def __init__(self, seed: int):
        self.state = seed or 0x9E3779B97F4A7C15
        self.spare = None

    def uniform(self) -> float:
        x = self.state
        x ^= x >> 12
        x ^= (x << 25) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 27
        self.state = x
        return (((x * 2685821657736338717) & 0xFFFFFFFFFFFFFFFF) >> 11) * 2.0**-53

    def normal(self) -> float:
        if self.spare is not None:
            out, self.spare = self.spare, None
            return out
        u1 = max(self.uniform(), np.finfo(np.float64).tiny)
        u2 = self.uniform()
        radius = np.sqrt(-2.0 * np.log(u1))
        angle = 6.283185307179586476925286766559 * u2
        self.spare = radius * np.sin(angle)
        return radius * np.cos(angle)


def paper_matrix(test_set: str, order: int | None = None) -> dict:
    """The symmetric paper test sets of the reference
    (preparePaperMatrix, src/accelerate_backend.cpp:690-801):

    * V / VI: a symmetric matrix with uniform [0, 1) (V) or standard normal
      (VI) entries drawn down each column from the diagonal (upper triangle
      mirrored), from the stream seeded 0x243f6a8885a308d3 ^
      (0x9e3779b97f4a7c15 * ordinal) (ordinal 5 or 6), then shifted by
      max(0, max(1, 0.01 ||A||_inf) - lambda_min) on the diagonal;
    * VII: the Hilbert matrix 1 / (i + j + 1) plus 1e-6 I.

    Returns A (column-major n x n as a numpy array), its condition number,
    the diagonal shift, the seed and the labels the reference records."""
    if test_set not in PAPER_ORDERS:
        raise ValueError("symmetric paper test sets are V, VI and VII")
    n = int(order or PAPER_ORDERS[test_set])
    out = {"paper_test_set": test_set, "order": n, "seed": 0}
    if test_set == "VII":
        i = np.arange(n, dtype=np.float64)
        A = 1.0 / (i[:, None] + i[None, :] + 1.0)
        A[np.diag_indices(n)] += 1.0e-6
        lam = np.linalg.eigvalsh(A)
        out.update(A=A, shift=1.0e-6, condition=lam[-1] / lam[0], family="hilbert_shifted",
                   shift_rule="fixed_hilbert_tau")
        return out
    seed = (0x243F6A8885A308D3 ^ ((0x9E3779B97F4A7C15 * _PAPER_ORDINALS[test_set])
                                  & 0xFFFFFFFFFFFFFFFF))
    st = _ScalarStream(seed)
    draw = st.uniform if test_set == "V" else st.normal
    A = np.empty((n, n))
    for j in range(n):
        col = np.array([draw() for _ in range(n - j)])
        A[j:, j] = col
        A[j, j:] = col
    separation = max(1.0, 0.01 * np.abs(A).sum(axis=1).max())
    lam = np.linalg.eigvalsh(A)
    shift = max(0.0, separation - lam[0])
    A[np.diag_indices(n)] += shift
    out.update(A=A, shift=shift, condition=(lam[-1] + shift) / (lam[0] + shift), seed=seed,
               family="uniform_random" if test_set == "V" else "normal_random",
               shift_rule="set_min_real_part_to_max_1_or_0.01_inf_norm")
    return out
