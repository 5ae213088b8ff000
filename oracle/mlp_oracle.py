"""CPU restatement (float64, numpy) of the warm-start network's loss and
gradient.  TEST INFRASTRUCTURE ONLY: the GPU training kernels are checked
against it; the product never imports it.

The network has no counterpart in the reference (it replaces the random
start vector of BLQH_without_spin/lanczos_hamiltonian.cpp:466-483).
Definition restated here:
  x   = (2 b_k - 1 for the nbits occupation bits b_k of the word, d / 3, 10 dsas)
  h1  = tanh(x W1 + b1),  h2 = tanh(h1 W2 + b2),  o = h2 W3 + b3   (64, 64, 2 units)
  L   = mean_s (o0_s - ya_s)^2 + softplus(o1_s) - ys_s o1_s
Flat parameters: W1[nin][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2],
nin = nbits + 2.  This restatement is pinned by central finite differences
(tests/test_mlp_oracle.py), not by retyping the kernels.
"""
import numpy as np

H = 64


def unpack(theta, nbits):
    nin = nbits + 2
    t = np.asarray(theta, np.float64)
    o = 0
    W1 = t[o:o + nin * H].reshape(nin, H); o += nin * H
    b1 = t[o:o + H]; o += H
    W2 = t[o:o + H * H].reshape(H, H); o += H * H
    b2 = t[o:o + H]; o += H
    W3 = t[o:o + 2 * H].reshape(H, 2); o += 2 * H
    b3 = t[o:o + 2]; o += 2
    assert o == t.size
    return W1, b1, W2, b2, W3, b3


def nparams(nbits):
    return (nbits + 2) * H + H + H * H + H + 2 * H + 2


def features(words, d, dsas, nbits):
    w = np.asarray(words, np.uint64)
    bits = ((w[:, None] >> np.arange(nbits, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float64)
    return np.concatenate([2.0 * bits - 1.0, np.asarray(d, np.float64)[:, None] / 3.0,
                           10.0 * np.asarray(dsas, np.float64)[:, None]], axis=1)


def forward(theta, nbits, X):
    W1, b1, W2, b2, W3, b3 = unpack(theta, nbits)
    h1 = np.tanh(X @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h1, h2, h2 @ W3 + b3


def loss(theta, nbits, words, d, dsas, ya, ys):
    X = features(words, d, dsas, nbits)
    _, _, o = forward(theta, nbits, X)
    ya = np.asarray(ya, np.float64)
    ys = np.asarray(ys, np.float64)
    return float(np.mean((o[:, 0] - ya) ** 2 + np.logaddexp(0.0, o[:, 1]) - ys * o[:, 1]))


def loss_grad(theta, nbits, words, d, dsas, ya, ys):
    """(mean loss, flat gradient) by back-propagation in float64."""
    X = features(words, d, dsas, nbits)
    W1, b1, W2, b2, W3, b3 = unpack(theta, nbits)
    h1, h2, o = forward(theta, nbits, X)
    ya = np.asarray(ya, np.float64)
    ys = np.asarray(ys, np.float64)
    B = X.shape[0]
    L = float(np.mean((o[:, 0] - ya) ** 2 + np.logaddexp(0.0, o[:, 1]) - ys * o[:, 1]))
    go = np.stack([2.0 * (o[:, 0] - ya), 1.0 / (1.0 + np.exp(-o[:, 1])) - ys], axis=1) / B
    gW3 = h2.T @ go
    gb3 = go.sum(0)
    d2 = (go @ W3.T) * (1.0 - h2 ** 2)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * (1.0 - h1 ** 2)
    gW1 = X.T @ d1
    gb1 = d1.sum(0)
    return L, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def start_vector(theta, nbits, words, d, dsas):
    """psi_i = exp(o0) sign(o1), normalised (real)."""
    n = len(words)
    X = features(words, np.full(n, d), np.full(n, dsas), nbits)
    _, _, o = forward(theta, nbits, X)
    v = np.exp(o[:, 0]) * np.where(o[:, 1] > 0, 1.0, -1.0)
    return v / np.linalg.norm(v)
