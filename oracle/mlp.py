"""CPU oracle of the learned tier selector (test infrastructure only).

The selector has no counterpart in the reference: it is the learned part of
the batched solver (per-matrix features, a 16-64-64-8 ReLU MLP, masked binary
cross-entropy, Adam).  This module restates the HIP library's definitions in
numpy so that the GPU kernels can be checked:

* ``features``  -- the 16 per-matrix features of the features kernel
                   (power iteration on A rounded to float32, in binary64:
                   16 steps on A from the normalised ones vector; 24 steps on
                   (c I - A), c = 1.0001 lambda_max, from a fixed hash start
                   vector; every other feature from A in binary64);
* ``forward``   -- logits and hidden activations;
* ``forward_bf16`` -- the BF16 inference kernel: weights, inputs and hidden
                   activations rounded to BF16 (nearest even), float32
                   accumulation in index order (each BF16 x BF16 product is
                   exact in float32, so a fused multiply-add is one rounded
                   float32 addition);
* ``backward``  -- the gradient of the masked BCE loss;
* ``adam``      -- one bias-corrected Adam step.

Pinning: ``backward`` is checked against central finite differences of
``loss`` in float64 (tests/test_selector_oracle.py); ``features`` against
closed forms (diagonal matrices, where the power iterations and Gershgorin
bounds are exact, and eigenvalue bounds on random SPD matrices).
"""
from __future__ import annotations

import numpy as np

INPUTS, HIDDEN, OUTPUTS = 16, 64, 8
POWER_MAX, POWER_MIN = 16, 24
SIZES = [("W1", (HIDDEN, INPUTS)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)),
         ("b2", (HIDDEN,)), ("W3", (OUTPUTS, HIDDEN)), ("b3", (OUTPUTS,))]
PARAMS = sum(int(np.prod(s)) for _, s in SIZES)  # 5768


def unpack(params):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + k]).reshape(shape)
        o += k
    return out


def pack(parts):
    return np.concatenate([np.asarray(parts[name]).reshape(-1) for name, _ in SIZES])


def start_vector(n):
    """Start of the lambda_min iteration: ((j + 1) * 2654435761 mod 2^32) >> 16,
    low 16 bits, scaled to [-0.5, 0.5)."""
    j = np.arange(n, dtype=np.uint64)
    h = ((j + 1) * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)
    return ((h >> np.uint64(16)) & np.uint64(0xFFFF)).astype(np.float64) / 65536.0 - 0.5


def _power(A, x, c, iters):
    last = prev = 0.0
    for _ in range(iters):
        y = A @ x if c == 0.0 else c * x - A @ x
        q = float(x @ y)
        yy = float(y @ y)
        prev, last = last, q
        x = y / np.sqrt(yy) if yy > 0 else y * 0.0
    return last, prev


def features(A):
    """The features kernel's 16 features of one SPD matrix (binary64)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    d = np.diag(A).copy()
    r = np.abs(A).sum(axis=1) - np.abs(d)
    fro = float(np.sum(A * A))
    off = fro - float(np.sum(d * d))
    tr = float(d.sum())
    t = tr / n
    A32 = A.astype(np.float32).astype(np.float64)
    lmax, lmax_prev = _power(A32, np.full(n, 1.0 / np.sqrt(n)), 0.0, POWER_MAX)
    c = 1.0001 * max(lmax, 1e-300)
    s = start_vector(n)
    mu, mu_prev = _power(A32, s / np.sqrt(s @ s), c, POWER_MIN)
    lmin = c - mu
    lmax_c = max(lmax, 1e-300)
    lmin_c = max(lmin, lmax_c * 1e-16)
    logd = np.log10(np.maximum(d, 1e-300))
    dom = np.clip(d / np.maximum(r, 1e-300), 1e-6, 1e6)
    f = np.empty(INPUTS)
    f[0] = n / 256.0
    f[1] = np.log2(n) / 8.0
    f[2] = np.log10(lmax_c / lmin_c)
    f[3] = np.log10(lmax_c / t)
    f[4] = np.log10(lmin_c / t)
    f[5] = min(max((d - r).min() / lmax_c, -16.0), 1.0)
    f[6] = np.log10(max((d + r).max(), 1e-300) / lmax_c)
    f[7] = np.log10(max(d.min(), 1e-300) / max(d.max(), 1e-300))
    f[8] = np.sqrt(max(np.mean(logd ** 2) - np.mean(logd) ** 2, 0.0))
    f[9] = np.log10(dom.min())
    f[10] = np.log10(dom.mean())
    f[11] = np.sqrt(max(off, 0.0) / fro) if fro > 0 else 0.0
    f[12] = np.log10(max(np.sqrt(fro) * np.sqrt(n) / tr, 1e-300))
    f[13] = np.log10(abs(lmax - lmax_prev) / lmax_c + 1e-30)
    f[14] = np.log10(abs(mu - mu_prev) / max(abs(mu), 1e-300) + 1e-30)
    f[15] = float(np.mean(d < r))
    return f


def forward(params, x):
    """Logits z and post-ReLU h1, h2 for inputs x (batch x 16)."""
    p = unpack(np.asarray(params, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ p["W1"].T + p["b1"], 0.0)
    h2 = np.maximum(h1 @ p["W2"].T + p["b2"], 0.0)
    z = h2 @ p["W3"].T + p["b3"]
    return z, h1, h2


def _bf16(v):
    """Round float32 values to BF16 (nearest even; NaN quieted) as float32."""
    b = np.asarray(v, np.float32).view(np.uint32).astype(np.uint64)
    r = b >> np.uint64(16)
    rem = b & np.uint64(0xFFFF)
    up = (rem > 0x8000) | ((rem == 0x8000) & ((r & np.uint64(1)) == 1))
    r = r + up.astype(np.uint64)
    nan = ((b & np.uint64(0x7F800000)) == 0x7F800000) & ((b & np.uint64(0x7FFFFF)) != 0)
    r = np.where(nan, (b >> np.uint64(16)) | np.uint64(0x40), r)
    return (r.astype(np.uint32) << np.uint32(16)).view(np.float32)


def forward_bf16(params, x):
    """Logits (float32) of the BF16 inference kernel for inputs x (batch x 16)."""
    p = {k: _bf16(np.asarray(v, np.float32)) for k, v in unpack(np.asarray(params, np.float32)).items()}
    x = _bf16(np.asarray(x, np.float32))
    f32 = np.float32

    def layer(inp, W, b):
        acc = np.broadcast_to(b, (inp.shape[0], b.size)).astype(f32).copy()
        for i in range(inp.shape[1]):
            acc = (acc + (inp[:, i:i + 1] * W[None, :, i]).astype(f32)).astype(f32)
        return acc

    h1 = _bf16(np.maximum(layer(x, p["W1"], p["b1"]), f32(0)))
    h2 = _bf16(np.maximum(layer(h1, p["W2"], p["b2"]), f32(0)))
    return layer(h2, p["W3"], p["b3"])


def loss(params, x, y, m):
    z, _, _ = forward(params, x)
    per = np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z))) - y * z
    return float(np.sum(m * per) / np.sum(m))


def backward(params, x, y, m):
    """Gradient of ``loss`` (masked BCE with logits, mean over the mask)."""
    p = unpack(np.asarray(params, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    z, h1, h2 = forward(params, x)
    dz = m * (1.0 / (1.0 + np.exp(-z)) - y) / np.sum(m)
    dh2 = (dz @ p["W3"]) * (h2 > 0)
    dh1 = (dh2 @ p["W2"]) * (h1 > 0)
    return pack({"W1": dh1.T @ x, "b1": dh1.sum(0), "W2": dh2.T @ h1, "b2": dh2.sum(0),
                 "W3": dz.T @ h2, "b3": dz.sum(0)})


def adam(params, grads, m1, m2, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, decay=0.0):
    """One Adam step; returns (params, m1, m2)."""
    g = grads + decay * params
    m1 = beta1 * m1 + (1 - beta1) * g
    m2 = beta2 * m2 + (1 - beta2) * g * g
    c1 = 1 - beta1 ** step
    c2 = 1 - beta2 ** step
    return params - lr * (m1 / c1) / (np.sqrt(m2 / c2) + eps), m1, m2
