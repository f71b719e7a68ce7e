"""Generates pow5_table.h: 128-bit normalised approximations of 5^q for
q in [-342, 308], used by the Eisel-Lemire decimal-to-binary step in cells.h
(Lemire, "Number Parsing at a Gigabyte per Second", 2021).

Entry T[q] lies in [2^127, 2^128) with 5^q ~= T * 2^(s - 127), s = floor(log2 5^q):
  q >= 0: the top 128 bits of 5^q (exact for q <= 55, truncated above);
  q <  0: ceil(2^(bitlen(5^-q) + 127) / 5^-q).
The script also checks that floor(q * log2 10) == (q * 217706) >> 16 over the
whole range, the integer form cells.h uses for the binary exponent.
"""
import math
import os

Q_MIN, Q_MAX = -342, 308


def entry(q):
    if q >= 0:
        b = 5 ** q
        n = b.bit_length()
        t = b << (128 - n) if n <= 128 else b >> (n - 128)
        s = n - 1
    else:
        b = 5 ** (-q)
        n = b.bit_length()
        t = -((-(1 << (n + 127))) // b)  # ceil
        s = -n
    assert (1 << 127) <= t < (1 << 128)
    return t, s


def main():
    lines = []
    for q in range(Q_MIN, Q_MAX + 1):
        t, s = entry(q)
        assert (q * 217706) >> 16 == s + q, q
        # floor(q log2 10) cross-check with exact integers
        if q >= 0:
            assert (1 << (s + q)) <= 10 ** q < (1 << (s + q + 1))
        lines.append("    {0x%016xULL, 0x%016xULL},  /* 5^%d */" % (t >> 64, t & ((1 << 64) - 1), q))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pow5_table.h")
    with open(out, "w") as f:
        f.write("// pow5_table.h -- generated by gen_pow5.py; do not edit.\n")
        f.write("// 128-bit normalised 5^q, q in [%d, %d]: {high, low}.\n" % (Q_MIN, Q_MAX))
        f.write("#pragma once\n#include <stdint.h>\n\n")
        f.write("#define FX_POW5_QMIN (%d)\n#define FX_POW5_QMAX (%d)\n" % (Q_MIN, Q_MAX))
        f.write("#define FX_POW5_TABLE_INIT { \\\n")
        f.write(" \\\n".join(lines))
        f.write(" \\\n}\n")
    print("wrote", out, Q_MAX - Q_MIN + 1, "entries")


if __name__ == "__main__":
    main()
