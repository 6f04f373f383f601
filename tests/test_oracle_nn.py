"""The network oracle (oracle/nn_oracle.py) pinned by mathematics: its
backward pass against central finite differences of its own forward pass in
exact FP64 mode, for wrap and zero padding, and its BF16 rounding against
hand-worked cases; its forward pass against explicit loops (convolution
taps, padding, pooling, upsampling and the whole network on an 8 x 8 image)."""
import numpy as np
import pytest

from oracle import nn_oracle as nno


def test_bf16_rounding_cases():
    one = 1.0
    ulp = 2.0 ** -7                       # BF16 spacing in [1, 2)
    cases = {one: one, one + ulp: one + ulp,
             one + ulp / 2: one,          # tie -> even (1.0)
             one + 1.5 * ulp: one + 2 * ulp,   # tie -> even (mantissa ...10)
             one + ulp / 2 + 2.0 ** -20: one + ulp,
             -3.0: -3.0, 0.0: 0.0}
    for a, b in cases.items():
        assert nno.bf16(np.array([a]))[0] == b, a


def test_layout_counts():
    offs, n = nno.layout()
    assert n == sum(co * 9 * ci + co for ci, co in zip(nno.CI, nno.CO))
    assert offs[0] == (0, 32 * 9)


@pytest.mark.parametrize("perx,pery,kind", [(True, True, 0), (False, True, 0), (False, False, 0),
                                            (True, True, 1), (False, True, 1)])
def test_backward_matches_finite_differences(perx, pery, kind):
    rng = np.random.default_rng(3)
    P = nno.init_params(seed=5)
    B, H, W = 2, 16, 16
    x = rng.standard_normal((B, H, W))
    t = rng.standard_normal((B, H, W))
    eg = (16 if perx else 15, 16 if pery else 13, perx, pery, 1.7)
    net = nno.Net(P, perx, pery, use_bf16=False)
    loss, g = net.loss_grad(x, t, kind, eg)
    offs, n = nno.layout()
    idx = []
    for wo, bo in offs:
        idx += list(rng.integers(wo, bo, 3)) + [bo]
    h = 1e-6
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[k] += h
        Pm[k] -= h
        lp, _ = nno.Net(Pp, perx, pery, use_bf16=False).loss_grad(x, t, kind, eg)
        lm, _ = nno.Net(Pm, perx, pery, use_bf16=False).loss_grad(x, t, kind, eg)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-7, (k, fd, g[k])


def test_padding_modes_differ_at_edges_only():
    """wrap and zero padding agree on interior pixels of the first conv"""
    P = nno.init_params(seed=1)
    x = np.random.default_rng(0).standard_normal((1, 16, 16))
    a = nno.Net(P, True, True, use_bf16=False)
    b = nno.Net(P, False, False, use_bf16=False)
    a.forward(x); b.forward(x)
    np.testing.assert_allclose(a.s["a0"][:, 1:-1, 1:-1], b.s["a0"][:, 1:-1, 1:-1], rtol=1e-12)
    assert not np.allclose(a.s["a0"][:, 0], b.s["a0"][:, 0])


def test_s1_is_the_5_point_operator():
    """S1 against a dense assembly of the grid-unit 5-point operator"""
    rng = np.random.default_rng(0)
    for mx, my, upx, upy in ((6, 5, False, False), (6, 4, True, False), (4, 6, True, True)):
        n = mx * my
        A = np.zeros((n, n))
        for j in range(my):
            for i in range(mx):
                k = i + j * mx
                A[k, k] = 2 + 2 * 1.7
                for di, dj, c in ((1, 0, 1.0), (-1, 0, 1.0), (0, 1, 1.7), (0, -1, 1.7)):
                    ii, jj = i + di, j + dj
                    if upx: ii %= mx
                    if upy: jj %= my
                    if 0 <= ii < mx and 0 <= jj < my:
                        A[k, ii + jj * mx] -= c
        v = rng.standard_normal((1, my + 2, mx + 3))
        v[:, my:, :] = 0; v[:, :, mx:] = 0
        got = nno.s1(v, mx, my, upx, upy, 1.7)[0, :my, :mx].ravel()
        np.testing.assert_allclose(got, A @ v[0, :my, :mx].ravel(), rtol=1e-12, atol=1e-12)


def brute_conv(a, W, perx, pery):
    """3x3 convolution by explicit pixel and tap loops: out[b, y, x, :] =
    sum over taps t of W[:, t, :] a[b, y + t//3 - 1, x + t%3 - 1, :], the
    source pixel wrapped on a periodic axis and zero outside otherwise"""
    B, H, Wd, _ = a.shape
    out = np.zeros((B, H, Wd, W.shape[0]))
    for b in range(B):
        for y in range(H):
            for x in range(Wd):
                for t in range(9):
                    sy, sx = y + t // 3 - 1, x + t % 3 - 1
                    if pery:
                        sy %= H
                    elif not 0 <= sy < H:
                        continue
                    if perx:
                        sx %= Wd
                    elif not 0 <= sx < Wd:
                        continue
                    out[b, y, x] += W[:, t, :] @ a[b, sy, sx]
    return out


@pytest.mark.parametrize("perx,pery", [(True, True), (False, True), (True, False), (False, False)])
def test_conv_against_brute_force(perx, pery):
    rng = np.random.default_rng(7)
    a = rng.standard_normal((2, 5, 6, 3))
    W = rng.standard_normal((4, 9, 3))
    net = nno.Net(nno.init_params(seed=0), perx, pery, use_bf16=False)
    np.testing.assert_allclose(net.conv(a, W), brute_conv(a, W, perx, pery), rtol=1e-12, atol=1e-12)


def test_single_tap_kernels_shift_the_image():
    """a kernel with one unit tap t reads the pixel at offset (t//3 - 1,
    t%3 - 1): a transposed or flipped tap order fails here"""
    a = np.arange(4 * 5, dtype=np.float64).reshape(1, 4, 5, 1)
    net = nno.Net(nno.init_params(seed=0), True, True, use_bf16=False)
    for t in range(9):
        W = np.zeros((1, 9, 1))
        W[0, t, 0] = 1.0
        dy, dx = t // 3 - 1, t % 3 - 1
        exp = np.roll(np.roll(a, -dy, axis=1), -dx, axis=2)
        np.testing.assert_array_equal(net.conv(a, W), exp)
        assert net.conv(a, W)[0, 1, 1, 0] == a[0, 1 + dy, 1 + dx, 0]


def test_pool_and_up_by_hand():
    a = np.arange(16, dtype=np.float64).reshape(1, 4, 4, 1)
    p = nno.Net.pool(a)
    assert p.shape == (1, 2, 2, 1)
    assert p[0, 0, 0, 0] == (0 + 1 + 4 + 5) / 4 and p[0, 1, 1, 0] == (10 + 11 + 14 + 15) / 4
    u = nno.Net.up(p)
    assert u.shape == (1, 4, 4, 1)
    np.testing.assert_array_equal(u[0, :, :, 0], np.kron(p[0, :, :, 0], np.ones((2, 2))))


@pytest.mark.parametrize("perx,pery", [(True, True), (False, False)])
def test_forward_against_loops(perx, pery):
    """the whole network on an 8 x 8 image, composed from the brute-force
    convolution and explicit 2 x 2 averages and copies"""
    rng = np.random.default_rng(11)
    P = nno.init_params(seed=3)
    x = rng.standard_normal((1, 8, 8))
    Ws, bs = nno.unpack(P)

    def layer(a, l, relu=True):
        o = brute_conv(a, Ws[l], perx, pery) + bs[l]
        return np.maximum(o, 0.0) if relu else o

    def pool(a):
        B, H, W, C = a.shape
        o = np.zeros((B, H // 2, W // 2, C))
        for y in range(H // 2):
            for xx in range(W // 2):
                o[:, y, xx] = (a[:, 2 * y, 2 * xx] + a[:, 2 * y, 2 * xx + 1] + a[:, 2 * y + 1, 2 * xx]
                               + a[:, 2 * y + 1, 2 * xx + 1]) / 4
        return o

    def up(a):
        B, H, W, C = a.shape
        o = np.zeros((B, 2 * H, 2 * W, C))
        for y in range(2 * H):
            for xx in range(2 * W):
                o[:, y, xx] = a[:, y // 2, xx // 2]
        return o

    a0 = layer(x[..., None], 0)
    a1 = layer(pool(a0), 1)
    a2 = layer(pool(a1), 2)
    a3 = layer(pool(a2), 3)
    u2 = up(layer(a3, 4)) + a2
    u1 = up(layer(u2, 5)) + a1
    u0 = up(layer(u1, 6)) + a0
    y = layer(u0, 7, relu=False)[..., 0]
    got = nno.Net(P, perx, pery, use_bf16=False).forward(x)
    np.testing.assert_allclose(got, y, rtol=1e-11, atol=1e-12)
