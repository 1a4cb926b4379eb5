"""CPU oracle of the learned-cutoff MLP (test infrastructure only; never shipped or timed).

Restates bosonsupremacy_amd/csrc/mlp.hip in NumPy: the in -> width -> width -> out ReLU
network, its gradients and the Adam step.  ``forward_bf16`` / ``backward_bf16`` round every
GEMM operand to BF16 (nearest-even) exactly as the GPU kernels do and accumulate in FP32, so
they differ from the GPU only by FP32 summation order.  ``forward64`` / ``backward64`` are the
same formulas in float64 without rounding; ``numeric_grad`` differentiates ``forward64`` by
central differences, which pins the analytic gradient formulas themselves.

The reference (sss441803/BosonSupremacy) has no learned component, so nothing here follows a
reference file; the formulas are standard backpropagation for this network.
"""
import numpy as np


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = rounded.astype(np.uint32).view(np.float32)
    return np.where(np.isfinite(x), out, x).astype(np.float32)


def forward_bf16(X, W1, b1, W2, b2, W3, b3):
    """(H1, H2, Y) with BF16 operands and FP32 accumulation; W* already BF16-valued."""
    f = np.float32
    H1 = np.maximum(bf16(X) @ W1.T.astype(f) + b1.astype(f), 0).astype(f)
    H2 = np.maximum(bf16(H1) @ W2.T.astype(f) + b2.astype(f), 0).astype(f)
    Y = (bf16(H2) @ W3.T.astype(f) + b3.astype(f)).astype(f)
    return H1, H2, Y


def backward_bf16(X, W2, W3, H1, H2, dY):
    """Gradients (dW1, db1, dW2, db2, dW3, db3) as the GPU computes them."""
    f = np.float32
    dW3 = (bf16(dY).T @ bf16(H2)).astype(f)
    db3 = dY.sum(axis=0, dtype=f)
    dZ2 = ((bf16(dY) @ W3.astype(f)) * (H2 > 0)).astype(f)
    dW2 = (bf16(dZ2).T @ bf16(H1)).astype(f)
    db2 = dZ2.sum(axis=0, dtype=f)
    dZ1 = ((bf16(dZ2) @ W2.astype(f)) * (H1 > 0)).astype(f)
    dW1 = (bf16(dZ1).T @ bf16(X)).astype(f)
    db1 = dZ1.sum(axis=0, dtype=f)
    return dW1, db1, dW2, db2, dW3, db3


def forward64(X, W1, b1, W2, b2, W3, b3):
    H1 = np.maximum(X @ W1.T + b1, 0)
    H2 = np.maximum(H1 @ W2.T + b2, 0)
    return H1, H2, H2 @ W3.T + b3


def backward64(X, W1, b1, W2, b2, W3, b3, dY):
    H1, H2, _ = forward64(X, W1, b1, W2, b2, W3, b3)
    dW3, db3 = dY.T @ H2, dY.sum(axis=0)
    dZ2 = (dY @ W3) * (H2 > 0)
    dW2, db2 = dZ2.T @ H1, dZ2.sum(axis=0)
    dZ1 = (dZ2 @ W2) * (H1 > 0)
    return dZ1.T @ X, dZ1.sum(axis=0), dW2, db2, dW3, db3


def mse(Y, T):
    return float(np.mean((Y - T) ** 2))


def numeric_grad(params, X, T, which, idx, h=1e-6):
    """Central difference of the float64 MSE loss with respect to params[which][idx]."""
    p = [np.array(q, dtype=np.float64, copy=True) for q in params]
    p[which][idx] += h
    lp = mse(forward64(X, *p)[2], T)
    p[which][idx] -= 2 * h
    lm = mse(forward64(X, *p)[2], T)
    return (lp - lm) / (2 * h)


def adam_step(p, g, m, v, lr, beta1, beta2, eps, step):
    """One Adam step in float32 (returns new p, m, v); the constants 1 - beta are formed in
    float32 as the kernel forms them."""
    f = np.float32
    b1, b2 = f(beta1), f(beta2)
    m = (b1 * m + (f(1) - b1) * g).astype(f)
    v = (b2 * v + (f(1) - b2) * g * g).astype(f)
    c1, c2 = f(1) - b1 ** f(step), f(1) - b2 ** f(step)
    p = (p - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))).astype(f)
    return p, m, v
