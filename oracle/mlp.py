"""Oracle of the learned warm start network (TEST INFRASTRUCTURE ONLY).

FP64 numpy restatement of the per-segment MLP the library trains on the GPU:
three dense layers nin -> H -> H -> 6, tanh on the hidden layers, weights
rounded to BF16 (round-to-nearest-even) for the forward pass, mean squared
error on standardised targets.  Parameters are packed as
[W1 (nin x H), b1, W2 (H x H), b2, W3 (H x 6), b3], W row-major inputs x
outputs.  The network has no counterpart in the reference (whose warm starts
are extrapolations, segment.cpp:53-73, filament.cpp:908-1016).
"""
import numpy as np

NIN, NOUT = 51, 6      # 27 state features + 4 lags x 6 previous targets


def bf16_round(w):
    """Round float32 values to bfloat16 (nearest, ties to even), as float64."""
    w = np.asarray(w, dtype=np.float32)
    u = w.view(np.uint32).astype(np.uint64)
    nan = np.isnan(w)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def unpack(p, nh, nin=NIN, nout=NOUT):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + nin * nh].reshape(nin, nh); o += nin * nh
    b1 = p[o:o + nh]; o += nh
    W2 = p[o:o + nh * nh].reshape(nh, nh); o += nh * nh
    b2 = p[o:o + nh]; o += nh
    W3 = p[o:o + nh * nout].reshape(nh, nout); o += nh * nout
    b3 = p[o:o + nout]; o += nout
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(p, X, nh, quantise=True):
    W1, b1, W2, b2, W3, b3 = unpack(p, nh)
    q = bf16_round if quantise else (lambda w: np.asarray(w, np.float64))
    X = np.asarray(X, np.float64)
    H1 = np.tanh(X @ q(W1) + b1)
    H2 = np.tanh(H1 @ q(W2) + b2)
    return H2 @ q(W3) + b3, (X, H1, H2)


def loss_and_grad(p, X, T, nh, quantise=True):
    """MSE over all B x 6 outputs and its gradient (through the rounded
    weights, i.e. the straight-through gradient w.r.t. the FP32 masters)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, nh)
    q = bf16_round if quantise else (lambda w: np.asarray(w, np.float64))
    Y, (X, H1, H2) = forward(p, X, nh, quantise)
    T = np.asarray(T, np.float64)
    n = Y.size
    loss = float(np.mean((Y - T) ** 2))
    dY = 2.0 * (Y - T) / n
    gW3 = H2.T @ dY
    gb3 = dY.sum(0)
    dH2 = (dY @ q(W3).T) * (1 - H2 ** 2)
    gW2 = H1.T @ dH2
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ q(W2).T) * (1 - H1 ** 2)
    gW1 = X.T @ dH1
    gb1 = dH1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
