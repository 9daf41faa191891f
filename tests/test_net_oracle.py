"""The network oracle (oracle/pynet.py) against finite differences (CPU)."""
import numpy as np

from oracle import pynet


def test_layer_table():
    names = [(n, ci, co, l) for n, ci, co, l in pynet.layer_shapes()]
    assert names == [("E0", 1, 16, 0), ("E1", 16, 32, 1), ("E2", 32, 64, 2), ("E3", 64, 64, 3),
                     ("D3", 64, 64, 3), ("D2", 64, 32, 2), ("D1", 32, 16, 1), ("D0", 16, 1, 0)]
    assert pynet.num_weights() == sum(9 * ci * co for _, ci, co, _ in pynet.layer_shapes())


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5, 6, 3))
    W = rng.normal(size=(9, 3, 2))
    y = pynet.conv(x, W)
    ref = np.zeros((5, 6, 2))
    for yy in range(5):
        for xx in range(6):
            for t in range(9):
                ky, kx = divmod(t, 3)
                sy, sx = yy + ky - 1, xx + kx - 1
                if 0 <= sy < 5 and 0 <= sx < 6:
                    ref[yy, xx] += x[sy, sx] @ W[t]
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_gradients_by_finite_differences():
    rng = np.random.default_rng(1)
    L, c0, cmax = 3, 4, 8          # a small network of the same shape family
    w = pynet.init_weights(L, c0, cmax, seed=2)
    w[-9 * c0:] = rng.normal(0, 0.3, 9 * c0)  # non-zero output layer
    x = rng.normal(size=(8, 8))
    gy = rng.normal(size=(8, 8))
    gw, gx = pynet.backward(w, x, gy, L, c0, cmax)
    f = lambda ww, xx: float(np.sum(gy * pynet.forward(ww, xx, L, c0, cmax)))
    eps = 1e-6
    for idx in rng.choice(w.size, 25, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd = (f(wp, x) - f(wm, x)) / (2 * eps)
        assert abs(fd - gw[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, gw[idx])
    for _ in range(10):
        i, j = rng.integers(0, 8, 2)
        xp, xm = x.copy(), x.copy()
        xp[i, j] += eps
        xm[i, j] -= eps
        fd = (f(w, xp) - f(w, xm)) / (2 * eps)
        assert abs(fd - gx[i, j]) <= 1e-6 * max(1.0, abs(fd))


def test_positive_homogeneity():
    rng = np.random.default_rng(3)
    w = pynet.init_weights(3, 4, 8, seed=4)
    w[-36:] = rng.normal(0, 0.3, 36)
    x = rng.normal(size=(8, 8))
    np.testing.assert_allclose(pynet.forward(w, 3.5 * x, 3, 4, 8),
                               3.5 * pynet.forward(w, x, 3, 4, 8), rtol=1e-12, atol=1e-12)


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -20, -2.5,
                  3.14159265, 0.0, -0.0])
    want = np.array([1.0, 1.0, 1 + 2.0 ** -6, 1 + 2.0 ** -7, -2.5, 3.140625, 0.0, -0.0])
    np.testing.assert_array_equal(pynet.bf16(v), want)


def test_bf16_rounding_matches_torch():
    torch = __import__("torch")
    rng = np.random.default_rng(3)
    v = (rng.normal(size=20000) * 10.0 ** rng.integers(-20, 20, 20000)).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(pynet.bf16(v), ref)


def test_bf16_model_is_close_to_exact():
    """The BF16 number model only perturbs the FP64 network slightly (forward)."""
    rng = np.random.default_rng(4)
    w = pynet.init_weights(seed=1)
    w[-9 * 16:] = rng.normal(0.0, 0.1, 9 * 16)
    x = rng.normal(size=(16, 32))
    y = pynet.forward(w, x)
    yb = pynet.forward(w, x, rnd=pynet.bf16)
    assert 0 < np.linalg.norm(yb - y) < 2e-2 * np.linalg.norm(y)


def _small(seed):
    rng = np.random.default_rng(seed)
    L, c0, cmax = 3, 4, 8
    w = pynet.init_weights(L, c0, cmax, seed=seed)
    w[-9 * c0:] = rng.normal(0, 0.3, 9 * c0)
    return rng, L, c0, cmax, w


def test_tied_network_is_symmetric_positive():
    """Tied weights: y = sum_l 4^l F_l^T F_l x, so (u, N v) = (N u, v) and
    (x, N x) >= 0 for any inputs (to rounding)."""
    rng, L, c0, cmax, w = _small(21)
    u, v = rng.normal(size=(8, 16)), rng.normal(size=(8, 16))
    Nu = pynet.forward(w, u, L, c0, cmax, tied=True)
    Nv = pynet.forward(w, v, L, c0, cmax, tied=True)
    assert abs(np.sum(u * Nv) - np.sum(Nu * v)) < 1e-10 * np.sum(np.abs(u * Nv))
    for k in range(5):
        x = rng.normal(size=(8, 16))
        assert np.sum(x * pynet.forward(w, x, L, c0, cmax, tied=True)) >= 0


def test_tied_network_equals_sum_of_squares():
    """With pools P and encoder maps F_l, the tied network is
    sum_l 4^l F_l^T F_l: checked by assembling the dense matrix."""
    rng, L, c0, cmax, w = _small(22)
    H, W = 8, 8
    n = H * W
    M = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        M[:, k] = pynet.forward(w, e.reshape(H, W), L, c0, cmax, tied=True).reshape(-1)
    np.testing.assert_allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > -1e-10 * np.abs(M).max()
    # energy of the encoder features: x^T M x = sum_l 4^l |F_l x|^2
    x = rng.normal(size=(H, W))
    _, A = pynet.forward(w, x, L, c0, cmax, keep=True, tied=True)
    energy = sum(4.0 ** l * np.sum(A["enc"][l] ** 2) for l in range(L))
    np.testing.assert_allclose(x.reshape(-1) @ M @ x.reshape(-1), energy, rtol=1e-10)


def test_tied_gradients_by_finite_differences():
    rng, L, c0, cmax, w = _small(23)
    x, gy = rng.normal(size=(8, 8)), rng.normal(size=(8, 8))
    gw, _ = pynet.backward(w, x, gy, L, c0, cmax, tied=True)
    f = lambda ww: float(np.sum(gy * pynet.forward(ww, x, L, c0, cmax, tied=True)))
    eps = 1e-6
    nE = sum(9 * ci * co for name, ci, co, _ in pynet.layer_shapes(L, c0, cmax) if name[0] == "E")
    for k in rng.choice(nE, 40, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (f(wp) - f(wm)) / (2 * eps)
        assert abs(fd - gw[k]) < 1e-6 * (1 + abs(fd)), (k, fd, gw[k])
    assert not np.any(gw[nE:])
