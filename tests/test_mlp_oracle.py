"""The warm-start network's CPU restatement (oracle/mlp_oracle.py), pinned by
central finite differences and a closed-form case, so that the GPU kernels
can be checked against it."""
import numpy as np

from oracle import mlp_oracle as M


def _data(nbits, n, seed):
    rng = np.random.default_rng(seed)
    words = rng.integers(0, 1 << nbits, size=n, dtype=np.uint64)
    d = rng.uniform(0, 3, n)
    dsas = rng.choice([0.0, 0.05, 0.1], n)
    ya = rng.normal(-4, 2, n)
    ys = rng.integers(0, 2, n).astype(np.float64)
    return words, d, dsas, ya, ys


def test_gradient_matches_central_differences():
    nbits = 6
    rng = np.random.default_rng(0)
    theta = rng.normal(0, 0.3, M.nparams(nbits))
    data = _data(nbits, 40, 1)
    L, g = M.loss_grad(theta, nbits, *data)
    assert L == M.loss(theta, nbits, *data)
    idx = rng.choice(theta.size, 120, replace=False)
    # every block of the flat layout is probed: W1, b1, W2, b2, W3, b3
    nin = nbits + 2
    idx = np.concatenate([idx, [0, nin * 64, nin * 64 + 64, nin * 64 + 64 + 4096, theta.size - 130, theta.size - 2,
                                theta.size - 1]])
    h = 1e-6
    for i in idx:
        tp = theta.copy(); tp[i] += h
        tm = theta.copy(); tm[i] -= h
        fd = (M.loss(tp, nbits, *data) - M.loss(tm, nbits, *data)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_zero_network_closed_form():
    """All parameters zero: o = 0, so L = mean(ya^2) + log 2 and the only
    non-zero gradients are those of b3: (-2 mean(ya), 1/2 - mean(ys))."""
    nbits = 10
    words, d, dsas, ya, ys = _data(nbits, 50, 2)
    L, g = M.loss_grad(np.zeros(M.nparams(nbits)), nbits, words, d, dsas, ya, ys)
    assert np.isclose(L, np.mean(ya ** 2) + np.log(2.0), rtol=1e-14)
    assert np.isclose(g[-2], -2 * np.mean(ya), rtol=1e-13)
    assert np.isclose(g[-1], 0.5 - np.mean(ys), rtol=1e-13, atol=1e-15)
    assert np.all(g[:-2] == 0)


def test_feature_encoding():
    X = M.features(np.array([0b1011], np.uint64), np.array([1.5]), np.array([0.05]), 4)
    assert X.tolist() == [[1.0, 1.0, -1.0, 1.0, 0.5, 0.5]]


def test_start_vector_is_signed_exp_of_output():
    nbits = 5
    theta = np.zeros(M.nparams(nbits))
    W1, b1, W2, b2, W3, b3 = M.unpack(theta, nbits)
    theta[-2] = 0.3   # b3[0]: constant log-amplitude
    theta[-1] = -1.0  # b3[1]: every sign negative
    v = M.start_vector(theta, nbits, np.arange(7, dtype=np.uint64), 1.0, 0.0)
    assert np.allclose(v, -np.ones(7) / np.sqrt(7))
