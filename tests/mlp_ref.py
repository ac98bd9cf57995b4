"""numpy restatement of the dispatcher network (F -> H -> H -> A, ReLU,
masked MSE) used to check the HIP forward/backward kernels.  It is itself
pinned by central finite differences in tests/test_mlp_ref.py."""
import numpy as np


def unpack(p, F, H, A):
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + A * H].reshape(A, H); o += A * H
    b3 = p[o:o + A]; o += A
    return W1, b1, W2, b2, W3, b3


def forward(p, X, F, H, A):
    W1, b1, W2, b2, W3, b3 = unpack(p, F, H, A)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3.T + b3, (h1, h2)


def loss_and_grad(p, X, T, M, F, H, A):
    W1, b1, W2, b2, W3, b3 = unpack(p, F, H, A)
    out, (h1, h2) = forward(p, X, F, H, A)
    w = 1.0 / max(1.0, float(M.sum()))
    e = (out - T) * M
    loss = float((e * e).sum() * w)
    dout = 2.0 * e * w
    gW3 = dout.T @ h2
    gb3 = dout.sum(0)
    dh2 = (dout @ W3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
