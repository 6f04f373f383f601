"""CPU oracle of the learned Poisson warm start / preconditioner network.

Test infrastructure only: the product network runs in gample_amd's HIP
kernels (gp_nn.hip); this module restates the same network in NumPy so the
tests can check the device forward pass and the device gradients.  There is
no reference counterpart (the reference solves with AMGCL only,
EMField.mm:866), so the restatement is pinned by mathematics instead: in
exact mode (FP64, no BF16 rounding) its backward pass is checked against
central finite differences of its forward pass.

Network (images NHWC, 3x3 convolutions, wrap padding on periodic axes and
zero padding otherwise; c0 = 32):
    a0 = relu(conv_in(x))          1 -> c0
    a1 = relu(W1 pool(a0))         c0 -> 2c0
    a2 = relu(W2 pool(a1))         2c0 -> 4c0
    a3 = relu(W3 pool(a2))         4c0 -> 4c0
    g2 = relu(W4 a3);  u2 = up(g2) + a2
    g1 = relu(W5 u2);  u1 = up(g1) + a1
    g0 = relu(W6 u1);  u0 = up(g0) + a0
    y  = conv_out(u0)              c0 -> 1
loss = mean((y - t)^2).  In BF16 mode every stored activation, every
BF16 weight copy (W1..W6) and every stored activation gradient is rounded to
BF16 (round to nearest even) at the points where the device stores BF16.
"""
import numpy as np

C0 = 32
CI = [1, C0, 2 * C0, 4 * C0, 4 * C0, 4 * C0, 2 * C0, C0]
CO = [C0, 2 * C0, 4 * C0, 4 * C0, 4 * C0, 2 * C0, C0, 1]


def layout():
    """(weight offset, bias offset) per layer in the flat parameter vector."""
    offs, o = [], 0
    for ci, co in zip(CI, CO):
        offs.append((o, o + co * 9 * ci))
        o += co * 9 * ci + co
    return offs, o


def unpack(P):
    offs, _ = layout()
    Ws, bs = [], []
    for (wo, bo), ci, co in zip(offs, CI, CO):
        Ws.append(np.asarray(P[wo:wo + co * 9 * ci], np.float64).reshape(co, 9, ci))
        bs.append(np.asarray(P[bo:bo + co], np.float64))
    return Ws, bs


def bf16(a):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


class Net:
    def __init__(self, P, perx, pery, use_bf16=True):
        self.Ws, self.bs = unpack(P)
        self.perx, self.pery = perx, pery
        self.q = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
        # BF16 copies of the MFMA layers' weights
        self.Wq = [W if l in (0, 7) else self.q(W) for l, W in enumerate(self.Ws)]

    # padded shift: out[b, y, x] = a[b, y + dy, x + dx] (wrap or zero)
    def shift(self, a, dy, dx):
        B, H, W = a.shape[:3]
        out = np.zeros_like(a)
        ys = np.arange(H) + dy
        xs = np.arange(W) + dx
        if self.pery:
            ys %= H
            vy = np.ones(H, bool)
        else:
            vy = (ys >= 0) & (ys < H)
        if self.perx:
            xs %= W
            vx = np.ones(W, bool)
        else:
            vx = (xs >= 0) & (xs < W)
        sub = a[:, ys[vy]][:, :, xs[vx]]
        out[:, np.nonzero(vy)[0][:, None], np.nonzero(vx)[0][None, :]] = sub
        return out

    def conv(self, a, W):
        """a [B,H,W,ci], W [co,9,ci] -> [B,H,W,co] (exact sums)"""
        out = 0.0
        for t in range(9):
            out = out + self.shift(a, t // 3 - 1, t % 3 - 1) @ W[:, t, :].T
        return out

    def conv_T(self, d, W):
        """adjoint of conv: d [B,H,W,co] -> [B,H,W,ci]"""
        out = 0.0
        for t in range(9):
            out = out + self.shift(d @ W[:, t, :], -(t // 3 - 1), -(t % 3 - 1))
        return out

    def conv_dw(self, d, a):
        """dW[co,t,ci] = sum_p d[p,co] a[p+off_t,ci]"""
        co, ci = d.shape[-1], a.shape[-1]
        dW = np.zeros((co, 9, ci))
        dd = d.reshape(-1, co)
        for t in range(9):
            dW[:, t, :] = dd.T @ self.shift(a, t // 3 - 1, t % 3 - 1).reshape(-1, ci)
        return dW

    @staticmethod
    def pool(a):
        return 0.25 * ((a[:, 0::2, 0::2] + a[:, 0::2, 1::2]) + (a[:, 1::2, 0::2] + a[:, 1::2, 1::2]))

    @staticmethod
    def up(a):
        return np.repeat(np.repeat(a, 2, axis=1), 2, axis=2)

    @staticmethod
    def sum2x2(a):
        return (a[:, 0::2, 0::2] + a[:, 0::2, 1::2]) + (a[:, 1::2, 0::2] + a[:, 1::2, 1::2])

    def forward(self, x):
        """x [B,H,W] -> y [B,H,W]; keeps the activations for backward"""
        q, Ws, bs, Wq = self.q, self.Ws, self.bs, self.Wq
        s = {}
        s["x"] = np.asarray(x, np.float64)[..., None]
        s["a0"] = q(np.maximum(self.conv(s["x"], Ws[0]) + bs[0], 0))
        s["p1"] = q(self.pool(s["a0"]))
        s["a1"] = q(np.maximum(self.conv(s["p1"], Wq[1]) + bs[1], 0))
        s["p2"] = q(self.pool(s["a1"]))
        s["a2"] = q(np.maximum(self.conv(s["p2"], Wq[2]) + bs[2], 0))
        s["p3"] = q(self.pool(s["a2"]))
        s["a3"] = q(np.maximum(self.conv(s["p3"], Wq[3]) + bs[3], 0))
        s["g2"] = q(np.maximum(self.conv(s["a3"], Wq[4]) + bs[4], 0))
        s["u2"] = q(self.up(s["g2"]) + s["a2"])
        s["g1"] = q(np.maximum(self.conv(s["u2"], Wq[5]) + bs[5], 0))
        s["u1"] = q(self.up(s["g1"]) + s["a1"])
        s["g0"] = q(np.maximum(self.conv(s["u1"], Wq[6]) + bs[6], 0))
        s["u0"] = q(self.up(s["g0"]) + s["a0"])
        s["y"] = (self.conv(s["u0"], Ws[7]) + bs[7])[..., 0]
        self.s = s
        return s["y"]

    def loss_grad(self, x, t, kind=0, egeo=None):
        """loss and the flat gradient (layout of unpack).  kind 0: mean
        square over the images; kind 1: energy norm d' S1 d per image (d = y - t
        on the unknown grid, zero in the padding), mean over the batch;
        egeo = (mx, my, periodic x, periodic y, cy/cx) of the unknown grid."""
        q, s = self.q, None
        y = self.forward(x)
        s = self.s
        diff = y - np.asarray(t, np.float64)
        if kind == 0:
            n = y.size
            loss = float(np.sum(diff * diff) / n)
            dy = (2.0 * diff / n)[..., None]
        else:
            mx, my, upx, upy, cyr = egeo
            d = np.zeros_like(diff)
            d[:, :my, :mx] = diff[:, :my, :mx]
            Sd = s1(d, mx, my, upx, upy, cyr)
            B = y.shape[0]
            loss = float(np.sum(d * Sd) / B)
            dy = (2.0 * Sd / B)[..., None]
        G = [None] * 8
        gb = [None] * 8
        G[7] = self.conv_dw(dy, s["u0"]); gb[7] = dy.sum(axis=(0, 1, 2))
        d_u0 = q(self.conv_T(dy, self.Ws[7]))
        dpre6 = q(np.where(s["g0"] > 0, self.sum2x2(d_u0), 0.0))
        G[6] = self.conv_dw(dpre6, s["u1"]); gb[6] = dpre6.sum(axis=(0, 1, 2))
        d_u1 = q(self.conv_T(dpre6, self.Wq[6]))
        dpre5 = q(np.where(s["g1"] > 0, self.sum2x2(d_u1), 0.0))
        G[5] = self.conv_dw(dpre5, s["u2"]); gb[5] = dpre5.sum(axis=(0, 1, 2))
        d_u2 = q(self.conv_T(dpre5, self.Wq[5]))
        dpre4 = q(np.where(s["g2"] > 0, self.sum2x2(d_u2), 0.0))
        G[4] = self.conv_dw(dpre4, s["a3"]); gb[4] = dpre4.sum(axis=(0, 1, 2))
        dpre3 = q(np.where(s["a3"] > 0, self.conv_T(dpre4, self.Wq[4]), 0.0))
        G[3] = self.conv_dw(dpre3, s["p3"]); gb[3] = dpre3.sum(axis=(0, 1, 2))
        d_p3 = q(self.conv_T(dpre3, self.Wq[3]))
        dpre2 = q(np.where(s["a2"] > 0, d_u2 + 0.25 * self.up(d_p3), 0.0))
        G[2] = self.conv_dw(dpre2, s["p2"]); gb[2] = dpre2.sum(axis=(0, 1, 2))
        d_p2 = q(self.conv_T(dpre2, self.Wq[2]))
        dpre1 = q(np.where(s["a1"] > 0, d_u1 + 0.25 * self.up(d_p2), 0.0))
        G[1] = self.conv_dw(dpre1, s["p1"]); gb[1] = dpre1.sum(axis=(0, 1, 2))
        d_p1 = q(self.conv_T(dpre1, self.Wq[1]))
        dpre0 = q(np.where(s["a0"] > 0, d_u0 + 0.25 * self.up(d_p1), 0.0))
        G[0] = self.conv_dw(dpre0, s["x"]); gb[0] = dpre0.sum(axis=(0, 1, 2))
        flat = np.concatenate([np.concatenate([G[l].ravel(), gb[l].ravel()]) for l in range(8)])
        return loss, flat


def s1(d, mx, my, upx, upy, cyr):
    """grid-unit Poisson operator on the unknown grid (mx x my) of padded
    images d [B, H, W]: (2u - uL - uR) + cyr (2u - uD - uU), wrap on a
    periodic axis, zero outside otherwise; zero in the padding"""
    u = d[:, :my, :mx]

    def nb(axis, step, per):
        if per:
            return np.roll(u, step, axis=axis)
        out = np.zeros_like(u)
        if axis == 2:
            if step == 1:
                out[:, :, 1:] = u[:, :, :-1]
            else:
                out[:, :, :-1] = u[:, :, 1:]
        else:
            if step == 1:
                out[:, 1:, :] = u[:, :-1, :]
            else:
                out[:, :-1, :] = u[:, 1:, :]
        return out

    out = np.zeros_like(d)
    out[:, :my, :mx] = ((2 * u - nb(2, 1, upx)) - nb(2, -1, upx)) + cyr * ((2 * u - nb(1, 1, upy)) - nb(1, -1, upy))
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    _, n = layout()
    P = np.zeros(n)
    offs, _ = layout()
    for l, ((wo, bo), ci, co) in enumerate(zip(offs, CI, CO)):
        lim = np.sqrt(6.0 / (9 * ci)) * (0.5 if l == 7 else 1.0)
        P[wo:bo] = rng.uniform(-lim, lim, co * 9 * ci)
        P[bo:bo + co] = rng.uniform(0.0, 0.1, co)
    return P
