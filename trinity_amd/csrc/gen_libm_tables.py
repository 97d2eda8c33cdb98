"""Writes trn_libm_tables.inc, the tables of trn_libm.hpp, from their
defining formulas (mpmath at 300 bits, then rounded to nearest double).

exp: 2^(k/128) ~= H[k] * (1 + T[k]) for k in [0, 128); stored as the pair
     (bits(T[k]), bits(H[k]) - (k << 52) / 128) so that adding k << 45 to
     the second word gives the scale of any exponent.
log: x = 2^k z with z in [0x1.69555p-1, 0x1.69555p0) split into 128
     sub-intervals by the top bits of (bits(z) - 0x3fe6955500000000); per
     sub-interval, with `center` its midpoint,
       1/c      = round(128/center)/128           (center < 1)
                = round(256/center)/256           (center >= 1)
       logc     = log(c) rounded to a multiple of 2^-43
       logctail = log(c) - logc rounded to double
     so that z/c - 1 is exact and k*ln2_hi + logc carries no rounding.

These are the tables of the exp/pow design published with glibc >= 2.28
(sysdeps/ieee754/dbl-64, from Arm's optimized-routines); regenerating them
from the formulas above reproduces glibc's values, which the CPU tests
confirm by comparing every function with the host libm.

    python trinity_amd/csrc/gen_libm_tables.py
"""
import os
import struct

import mpmath

mpmath.mp.prec = 300
N = 128
OFF = 0x3FE6955500000000


def _bits(d: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def _dbl(u: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", u))[0]


def exp_table():
    out = []
    for k in range(N):
        v = mpmath.mpf(2) ** (mpmath.mpf(k) / N)
        h = float(v)
        t = float((v - h) / h)
        out.append((_bits(t), (_bits(h) - ((k << 52) // N)) & 0xFFFFFFFFFFFFFFFF))
    return out


def log_table():
    rows = []
    for i in range(N):
        lo = _dbl(OFF + (i << 45))
        hi = _dbl(OFF + ((i + 1) << 45) - 1)
        center = (mpmath.mpf(lo) + mpmath.mpf(hi)) / 2
        if center < 1:
            invc = mpmath.mpf(int(mpmath.nint(N / center))) / N
        else:
            invc = mpmath.mpf(int(mpmath.nint(2 * N / center))) / N / 2
        logc_exact = -mpmath.log(invc)
        logc = mpmath.nint(logc_exact * mpmath.mpf(2) ** 43) / mpmath.mpf(2) ** 43
        rows.append((float(invc), float(logc), float(logc_exact - logc)))
    return rows


def main() -> None:
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "trn_libm_tables.inc")
    with open(path, "w") as f:
        f.write("// Generated by gen_libm_tables.py -- do not edit.\n")
        f.write("// exp: {bits(T[k]), bits(H[k]) - (k << 45)}, 2^(k/128) ~= H[k] (1 + T[k])\n")
        f.write("constexpr uint64_t kExpTab[2 * 128] = {\n")
        for t, h in exp_table():
            f.write(f"    0x{t:016x}ULL, 0x{h:016x}ULL,\n")
        f.write("};\n")
        f.write("// log: {1/c, logc, logctail} per sub-interval of [0x1.69555p-1, 0x1.69555p0)\n")
        f.write("constexpr double kLogTab[128][3] = {\n")
        for a, b, c in log_table():
            f.write(f"    {{{a.hex()}, {b.hex()}, {c.hex()}}},\n")
        f.write("};\n")


if __name__ == "__main__":
    main()
