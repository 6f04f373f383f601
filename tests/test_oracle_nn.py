"""The network oracle (oracle/nn_oracle.py) pinned by mathematics: its
backward pass against central finite differences of its own forward pass in
exact FP64 mode, for wrap and zero padding, and its BF16 rounding against
hand-worked cases."""
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
