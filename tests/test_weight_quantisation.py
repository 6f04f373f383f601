"""The device deposit quantises each per-axis weight s in [0, 1] to
rint(s * 2^23) with one fma, s * 2^23 + 2^23, keeping the low 24 bits of
the float encoding; the oracle uses rintf(s * 2^23) directly.  This checks
the identity for every float32 in [0, 1] (s * 2^23 is exact, so the fma's one
rounding is the float32 addition below)."""
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
