"""The device deposit quantises each per-axis weight s in [0, 1] to
rint(s * 2^23) with one fma, s * 2^23 + 2^23, keeping the low 24 bits of
the float encoding (orders 1 and 5), and to rint(s * 2^16) with the fma
s * 2^16 + 2^23, keeping the low 23 bits (order 3); the oracle uses
rintf(s * 2^k) directly.  These check the identities for every float32 in
[0, 1] (s * 2^k is exact, so the fma's one rounding is the float32 addition
below), and that no order-3 weight product needs more than 31 bits."""
import numpy as np


def test_q23_identity_exhaustive():
    one = np.float32(1.0).view(np.uint32)
    two23 = np.float32(2.0 ** 23)
    chunk = 1 << 25
    for start in range(0, int(one) + 1, chunk):
        bits = np.arange(start, min(start + chunk, int(one) + 1), dtype=np.uint32)
        s = bits.view(np.float32)
        prod = s * two23                                   # exact
        fast = (prod + two23).view(np.uint32) & np.uint32(0xFFFFFF)
        ref = np.rint(prod).astype(np.uint32)             # round half to even
        bad = np.nonzero(fast != ref)[0]
        assert bad.size == 0, (s[bad[:4]], fast[bad[:4]], ref[bad[:4]])


def test_q16_identity_exhaustive():
    one = np.float32(1.0).view(np.uint32)
    two16, two23 = np.float32(2.0 ** 16), np.float32(2.0 ** 23)
    chunk = 1 << 25
    for start in range(0, int(one) + 1, chunk):
        bits = np.arange(start, min(start + chunk, int(one) + 1), dtype=np.uint32)
        s = bits.view(np.float32)
        prod = s * two16                                   # exact
        fast = (prod + two23).view(np.uint32) & np.uint32(0x7FFFFF)
        ref = np.rint(prod).astype(np.uint32)
        bad = np.nonzero(fast != ref)[0]
        assert bad.size == 0, (s[bad[:4]], fast[bad[:4]], ref[bad[:4]])


def test_order3_products_fit_31_bits(orc):
    """the largest order-3 weight is w1(0) = w2(1) = 2/3: its 16-bit value and
    any product of two stay below 2^31 (one 24-bit multiply, 32-bit window
    words with a carry)"""
    h = np.linspace(0.0, 1.0, 20001, dtype=np.float32)
    wmax = max(max(orc.weights(3, float(v), "fast")) for v in h)
    q = int(np.rint(np.float32(wmax) * np.float32(2 ** 16)))
    assert q == 43691 and q * q < 2 ** 31
