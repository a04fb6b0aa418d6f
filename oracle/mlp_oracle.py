"""CPU oracle of the learned-priority MLP (TEST INFRASTRUCTURE ONLY).

Forward pass and mean-squared-error gradient of the 16 -> 64 -> 64 -> 1 ReLU
network, in NumPy float64 with the same BF16 roundings the GPU kernels apply
to MFMA operands (inputs, weights W1/W2/W3, hidden activations, and the
back-propagated dH2/dH1), so the device gradients can be checked against it.
Parameter layout: W1[16][64] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3.
"""
import numpy as np

IN, H = 16, 64


ROUND = True  # apply the kernels' BF16 operand roundings (tests may disable)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    if not ROUND:
        return np.asarray(x, np.float64)
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float32 if ROUND else np.float64)
    o = 0
    W1 = p[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(X, p):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Xb = bf16(X)
    H1 = np.maximum(Xb @ bf16(W1) + b1, 0.0)
    H2 = np.maximum(bf16(H1) @ bf16(W2) + b2, 0.0)
    y = H2 @ bf16(W3) + b3
    return y, (Xb, H1, H2)


def grad(X, t, p):
    """Gradient of mean((y - t)^2) with the kernels' operand roundings."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = len(X)
    y, (Xb, H1, H2) = forward(X, p)
    e = y - np.asarray(t, np.float64)
    dy = 2.0 * e / n
    gW3 = H2.T @ dy
    gb3 = dy.sum()
    dH2 = np.outer(dy, bf16(W3)) * (H2 > 0)
    gW2 = bf16(H1).T @ bf16(dH2)
    gb2 = dH2.sum(0)
    dH1 = (bf16(dH2) @ bf16(W2).T) * (H1 > 0)
    gW1 = Xb.T @ bf16(dH1)
    gb1 = dH1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return g, float(np.mean(e * e))
