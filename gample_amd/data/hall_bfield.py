"""Writes the 1D magnetic-field files of the Hall-thruster condition
(org_hall.json) in the format EMField's From1dXFile input reads
(EMField.mm:175-197; the writer is tools/makeBfld.py): int32 n, float32 dx,
then n + 1 float32 values in tesla, for Bx (nx + 1 values), By and Bz
(nx + 2 values), nx = ngx + 2 ngb.

Bz is the radial-field profile of a Hall thruster channel: a Gaussian of
width 0.625 cm that rises from 6 mT at the anode (x = 0) to 10 mT at
x = 0.75 cm and falls to 1 mT at the far wall, sampled at the cell-centred
positions (i + 1/2) dx of the Bz array; Bx and By are zero.

    python gample_amd/data/hall_bfield.py [ngx ngb dx out_prefix]
"""
import math
import sys

import numpy as np


def bz_profile(x, x_lo, x_hi, b_lo=6e-3, b_max=10e-3, b_hi=1e-3, x_max=0.75, sigma=0.625):
    """two Gaussians of width sigma joined at their common peak b_max (x_max),
    through b_lo at x_lo and b_hi at x_hi"""
    x_peak = x_lo + x_max
    def side(b_end, x_end):
        g = math.exp(-0.5 * (x_peak - x_end) ** 2 / sigma ** 2)
        a = (b_max - b_end) / (1.0 - g)
        return a, b_max - a
    a1, c1 = side(b_lo, x_lo)
    a2, c2 = side(b_hi, x_hi)
    g = np.exp(-0.5 * (x - x_peak) ** 2 / sigma ** 2)
    return np.where(x < x_peak, a1 * g + c1, a2 * g + c2)


def write(path, dx, values):
    v = np.asarray(values, np.float32)
    with open(path, "wb") as f:
        f.write(np.int32(len(v) - 1).tobytes())
        f.write(np.float32(dx).tobytes())
        f.write(v.tobytes())


def main(argv):
    ngx = int(argv[0]) if len(argv) > 0 else 500
    ngb = int(argv[1]) if len(argv) > 1 else 0
    dx = float(argv[2]) if len(argv) > 2 else 5e-3
    prefix = argv[3] if len(argv) > 3 else "gample_amd/data/hall_"
    nx = ngx + 2 * ngb
    x = (np.arange(nx + 2) + 0.5) * dx
    bz = bz_profile(x, ngb * dx, (ngx + ngb) * dx)
    write(prefix + "Bz.bin", dx, bz)
    write(prefix + "Bx.bin", dx, np.zeros(nx + 1))
    write(prefix + "By.bin", dx, np.zeros(nx + 2))
    print(f"wrote {prefix}B[xyz].bin: nx = {nx}, Bz {bz.min() * 1e3:.2f}..{bz.max() * 1e3:.2f} mT")


if __name__ == "__main__":
    main(sys.argv[1:])
