"""Pins of the warm-start network's CPU specification (oracle/mlp_oracle.py)."""
import numpy as np

from oracle import mlp_oracle as M


def batch(seed, n=24):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 1, (n, 18)).astype(np.float32)
    t = rng.normal(0, 1, (n, 3)).astype(np.float32)
    return f, t


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-9, -2.5, 1.0 + 2.0**-7, 1.0 + 2.0**-7 + 2.0**-8], np.float32)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-9 rounds up to 1+2^-7; 1+2^-7 is exact;
    # 1+2^-7+2^-8 is a tie -> even (1+2^-6)
    np.testing.assert_array_equal(M.bf16(x), np.array([1.0, 1.0, 1.0 + 2.0**-7, -2.5, 1.0 + 2.0**-7, 1.0 + 2.0**-6],
                                                       np.float32))
    assert (M.bf16(x).view(np.uint32) & 0xFFFF).max() == 0


def test_parameter_count_and_packing():
    assert M.NPARAM == 18 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 3 + 3 == 9731
    p = np.arange(M.NPARAM, dtype=np.float32)
    Ws, bs = M.unpack(p)
    assert Ws[0].shape == (18, 64) and Ws[0][1, 0] == 64 and bs[3].shape == (3,) and bs[3][-1] == 9730


def test_fp64_gradient_matches_central_differences():
    p = M.init_params(3).astype(np.float64)
    p += np.random.default_rng(4).normal(0, 0.05, p.shape)  # nonzero biases
    scales = np.array([0.5, 1.0, 2.0, 0.3, 0.7, 1.5, 0.8])
    f, t = batch(5)
    loss, g = M.loss_grad(p, scales, f, t, rounding=False)
    rng = np.random.default_rng(6)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    for i in idx:
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += e
        pm[i] -= e
        fd = (M.loss_grad(pp, scales, f, t, rounding=False)[0] - M.loss_grad(pm, scales, f, t, rounding=False)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_variant_tracks_fp64():
    p = M.init_params(7)
    scales = np.array([0.5, 1.0, 2.0, 0.3, 0.7, 1.5, 0.8], np.float32)
    f, t = batch(8, 64)
    o32, _ = M.forward(p, scales, f, rounding=True)
    o64, _ = M.forward(p, scales, f, rounding=False)
    assert np.abs(o32 - o64).max() <= 0.03 * np.abs(o64).max()
    l32, g32 = M.loss_grad(p, scales, f, t, rounding=True)
    l64, g64 = M.loss_grad(p, scales, f, t, rounding=False)
    assert abs(l32 - l64) <= 0.02 * l64
    assert np.linalg.norm(g32 - g64) <= 0.05 * np.linalg.norm(g64)


def test_features_translation_and_ring_means():
    # 3 nodes in a line, node 1 adjacent to 0 and 2
    rest = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], float)
    x = rest + np.array([[0, 0, 0], [0, 1, 0], [0, 3, 0]], float)
    v = np.zeros((3, 3))
    off = np.array([0, 1, 3, 4])
    adj = np.array([1, 0, 2, 1])
    f = M.features(x, v, rest, 0.1, (0, 0, -10.0), off, adj, [0, 0, 0])
    np.testing.assert_allclose(f[1, 0:3], [0, 1, 0])
    np.testing.assert_allclose(f[1, 9:12], [0, 0.5, 0])  # mean(0, 3) - 1
    np.testing.assert_allclose(f[0, 6:9], [0, 0, -0.1], rtol=1e-6)  # h^2 g
    f2 = M.features(x + 5.0, v, rest + 5.0, 0.1, (0, 0, -10.0), off, adj, [0, 0, 0])
    np.testing.assert_allclose(f, f2, atol=1e-6)
    assert not M.features(x, v, rest, 0.1, (0, 0, -10.0), off, adj, [1, 0, 0])[0].any()
