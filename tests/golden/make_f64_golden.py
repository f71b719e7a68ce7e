"""Writes tests/golden/f64_cells.json: expected `str::parse::<f64>` results for
float cells, computed with Python's float() (a correctly rounded decimal to
binary conversion, David Gay's algorithm) behind a regex of Rust's float
grammar (`[+-]?(inf|infinity|nan|(d+.?d*|.d+)([eE][+-]?d+)?)`, the words
case-insensitive). Independent of the oracle: it pins the oracle's parser.

Cells: the reference's hand cases (tests/columns.rs:408-451), IEEE edges,
round-half-even boundaries, long mantissas and random shapes.
"""
import json
import os
import random
import re
import struct

GRAMMAR = re.compile(r"^[+-]?(?:(?i:inf|infinity|nan)|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)$")
FIXED6 = re.compile(r"^-?\d{1,3}\.\d{6}$")


def bits(v: float) -> str:
    return "%016x" % struct.unpack("<Q", struct.pack("<d", v))[0]


def expected(cell: str):
    if not GRAMMAR.match(cell):
        return [cell, False, None, False]
    return [cell, True, bits(float(cell)), bool(FIXED6.match(cell))]


def cells():
    hand = ["1.5", "-0.0", "0.1", "0.0", "1_000", "1.5e308", "5e-324", "NaN", "inf", "-inf", "", "nan",
            "+nan", "-nan", "Infinity", "-INFINITY", "infinit", "1.", ".5", ".", "1e", "1e+", "e5", "+",
            "-", " 1", "1 ", "0x10", "1.5E+3", "1.5e-3", "+.5", "-.5e1", "00001", "1e99999", "1e-99999"]
    edges = ["2.2250738585072011e-308", "2.2250738585072014e-308", "2.225073858507201e-308",
             "4.9406564584124654e-324", "2.4703282292062327e-324", "2.4703282292062328e-324",
             "1.7976931348623157e308", "1.7976931348623158e308", "1.7976931348623159e308",
             "9007199254740992", "9007199254740993", "9007199254740994", "9007199254740995",
             "9007199254740992.5", "0.30000000000000004", "123456789012345678901234567890",
             "1e23", "8.98846567431158e307", "4.35e-07", "7.038531e-26", "2.470328229206232720882e-324"]
    # Exact halfway points between consecutive doubles, and one ulp-digit either side.
    rnd = random.Random(20261203)
    halfway = []
    for _ in range(60):
        m = rnd.getrandbits(53) | (1 << 52)
        e = rnd.randint(-1074 + 60, 970)
        from fractions import Fraction
        mid = Fraction(2 * m + 1) * Fraction(2) ** (e - 1)
        # exact decimal expansion of mid
        num, den = mid.numerator, mid.denominator
        k = 0
        while den % 10 != 0 and den != 1 and k < 1200:
            if den % 2 == 0 or den % 5 == 0:
                num *= 10
                k += 1
                g = __import__("math").gcd(num, den)
                num //= g
                den //= g
            else:
                break
        if den == 1:
            s = str(num)
            if k:
                s = s.rjust(k + 1, "0")
                s = s[:-k] + "." + s[-k:]
            halfway.append(s)
            halfway.append(s + "0000000000001")
            if "." in s and len(s) > 3:
                halfway.append(s[:-1] + str(max(0, int(s[-1]) - 1)))
    rand = []
    for _ in range(400):
        kind = rnd.randrange(5)
        sgn = rnd.choice(["", "-", "+"])
        if kind == 0:
            rand.append(f"{sgn}{rnd.randrange(1000)}.{rnd.randrange(10**6):06d}")
        elif kind == 1:
            rand.append(f"{sgn}{rnd.randrange(10**rnd.randint(1, 20))}e{rnd.randint(-340, 310)}")
        elif kind == 2:
            d = "".join(rnd.choice("0123456789") for _ in range(rnd.randint(16, 40)))
            rand.append(f"{sgn}{d[:rnd.randint(1, len(d))]}.{d}")
        elif kind == 3:
            rand.append(f"{sgn}.{rnd.randrange(10**15)}e-{rnd.randint(300, 330)}")
        else:
            rand.append(f"{sgn}{rnd.random() * 10**rnd.randint(-5, 5)!r}")
    return hand + edges + halfway + rand


def main():
    out = [expected(c) for c in cells()]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "f64_cells.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_f64_golden.py", "cells": out}, f, indent=0)
    print(f"wrote {len(out)} cells to {path}")


if __name__ == "__main__":
    main()
