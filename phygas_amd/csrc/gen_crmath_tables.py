"""Generate crmath_tables.h: the double-double constants of the correctly rounded pow / sin / cos (crmath.h).

    python phygas_amd/csrc/gen_crmath_tables.py          (rewrites crmath_tables.h beside this file)

The tables are namespace-scope constexpr arrays: host code and device code (where clang places a copy in
constant memory) index the same values.  Every value is computed with mpmath at 300 bits and split exactly into hi + lo doubles (hi = the nearest
double, lo = the nearest double to the remainder), through exact rationals so no conversion rounds twice.

  log table   i = 0..127: c_i = 1 / (1 + (i + 1/2) / 128) rounded to double, and -log(c_i) as hi + lo
  exp table   j = 0..63:  2^(j/64) as hi + lo
  sin/cos     j = 0..52:  sin(j/64), cos(j/64) as hi + lo
  constants   ln 2 with a 42-bit head (e * head exact for |e| < 2^11), ln2/64 with a 36-bit head (k * head exact for |k| < 2^17), 64/ln 2,
              pi/2 in three 33-bit pieces (k * piece exact for |k| < 2^20), 2/pi
"""
from __future__ import annotations

import os
from fractions import Fraction

import mpmath as mp

mp.mp.prec = 300
HERE = os.path.dirname(os.path.abspath(__file__))


def exact(x) -> Fraction:
    """The exact rational value of an mpf (or int / Fraction)."""
    if isinstance(x, (int, Fraction)):
        return Fraction(x)
    sign, man, exp, _ = mp.mpf(x)._mpf_
    v = Fraction(int(man)) * (Fraction(2) ** exp)
    return -v if sign else v


def split(x):
    """x = hi + lo: hi the double nearest x, lo the double nearest x - hi."""
    fx = exact(x)
    hi = float(fx)                 # Fraction -> float rounds to nearest
    lo = float(fx - Fraction(hi))
    return hi, lo


def head(x, frac_bits: int) -> float:
    """x rounded to a multiple of 2^-frac_bits (a short head, so small-integer multiples are exact)."""
    fx = exact(x) * (2 ** frac_bits)
    n = round(fx)
    return float(Fraction(n, 2 ** frac_bits))


def lit(v: float) -> str:
    return float.hex(v)


def main() -> None:
    ln2 = mp.log(2)
    out = ["// Generated by gen_crmath_tables.py (mpmath, 300 bits); do not edit.",
           "// Double-double constants of the correctly rounded pow, sin and cos in crmath.h.",
           "#pragma once", "", "namespace phx {", "namespace cr {", ""]
    ln2_hi = head(ln2, 42)
    ln2_lo = float(exact(ln2) - Fraction(ln2_hi))
    c1 = head(ln2 / 64, 42)
    c2 = float(exact(ln2 / 64) - Fraction(c1))
    inv = float(exact(64 / ln2))
    pio2 = mp.pi / 2
    p1 = head(pio2, 32)
    p2 = head(exact(pio2) - Fraction(p1), 65)
    p3 = float(exact(pio2) - Fraction(p1) - Fraction(p2))
    out += [f"constexpr double LN2_HI = {lit(ln2_hi)}, LN2_LO = {lit(ln2_lo)};",
            f"constexpr double LN2_64_HI = {lit(c1)}, LN2_64_LO = {lit(c2)}, INV_LN2_64 = {lit(inv)};",
            f"constexpr double PIO2_1 = {lit(p1)}, PIO2_2 = {lit(p2)}, PIO2_3 = {lit(p3)};",
            f"constexpr double TWO_OVER_PI = {lit(float(exact(2 / mp.pi)))};", ""]
    # log: c_i and -log(c_i)
    rows = []
    for i in range(128):
        m = Fraction(1) + Fraction(2 * i + 1, 256)
        c = float(Fraction(1) / m)
        hi, lo = split(-mp.log(mp.mpf(exact(c).numerator) / exact(c).denominator))
        rows.append(f"    {{{lit(c)}, {lit(hi)}, {lit(lo)}}},")
    out += ["// {c_i, -log(c_i) hi, lo}", "constexpr double LOG_TAB[128][3] = {"] + rows + ["};", ""]
    rows = []
    for j in range(64):
        hi, lo = split(mp.power(2, mp.mpf(j) / 64))
        rows.append(f"    {{{lit(hi)}, {lit(lo)}}},")
    out += ["// {2^(j/64) hi, lo}", "constexpr double EXP_TAB[64][2] = {"] + rows + ["};", ""]
    rows = []
    for j in range(53):
        a = mp.mpf(j) / 64
        sh, sl = split(mp.sin(a))
        ch, cl = split(mp.cos(a))
        rows.append(f"    {{{lit(sh)}, {lit(sl)}, {lit(ch)}, {lit(cl)}}},")
    out += ["// {sin(j/64) hi, lo, cos(j/64) hi, lo}", "constexpr double SC_TAB[53][4] = {"] + rows + ["};", ""]
    out += ["}  // namespace cr", "}  // namespace phx", ""]
    with open(os.path.join(HERE, "crmath_tables.h"), "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
