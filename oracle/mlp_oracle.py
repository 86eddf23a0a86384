"""Double-precision restatement of the selector's MLP (forward, masked MSE
loss, backprop) -- the checker for the HIP kernels in numc_amd/csrc/mlp.hip.

TEST INFRASTRUCTURE ONLY.  Parameter layout: W0[din][H], b0[H], W1[H][H],
b1[H], W2[H][dout], b2[dout]; ReLU after the first two layers;
loss = sum_i sum_j 0.5 * mask_ij * (y_ij - t_ij)^2 / B.
Pinned in tests/test_mlp.py against central finite differences.
"""
import numpy as np


def unpack(p, din, H, dout):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((din, H), (H,), (H, H), (H,), (H, dout), (dout,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size
    return out


def forward(p, x, din, H, dout):
    W0, b0, W1, b1, W2, b2 = unpack(p, din, H, dout)
    x = np.asarray(x, np.float64)
    h1 = np.maximum(x @ W0 + b0, 0)
    h2 = np.maximum(h1 @ W1 + b1, 0)
    return h2 @ W2 + b2, (x, h1, h2)


def loss_grad(p, x, t, mask, din, H, dout):
    W0, b0, W1, b1, W2, b2 = unpack(p, din, H, dout)
    y, (x, h1, h2) = forward(p, x, din, H, dout)
    B = x.shape[0]
    m = np.ones_like(y) if mask is None else np.asarray(mask, np.float64)
    d = (y - np.asarray(t, np.float64)) * m
    loss = 0.5 * np.sum(d * d) / B
    dy = d / B
    gW2 = h2.T @ dy
    gb2 = dy.sum(0)
    dh2 = (dy @ W2.T) * (h2 > 0)
    gW1 = h1.T @ dh2
    gb1 = dh2.sum(0)
    dh1 = (dh2 @ W1.T) * (h1 > 0)
    gW0 = x.T @ dh1
    gb0 = dh1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW0, gb0, gW1, gb1, gW2, gb2)])
    return loss, g
