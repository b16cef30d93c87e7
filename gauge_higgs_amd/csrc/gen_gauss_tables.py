"""Writes gauss_tables.hpp: the constants of the correctly rounded log and cos
of gauss_cr.hpp, each split into doubles (hi + lo [+ lo2]) from a 60-digit
decimal evaluation.

    python gauge_higgs_amd/csrc/gen_gauss_tables.py > gauge_higgs_amd/csrc/gauss_tables.hpp

- ln 2 and pi/2 in three parts;
- log(j/128) for j = 96..192 (the log table, m in [0.75, 1.5));
- sin(j/64) and cos(j/64) for j = 0..51 (the sin/cos table, |r| <= pi/4 + 1/128);
- the double-double series coefficients 1/3, 1/5, 1/6, 1/24, 1/120.

Only the Python standard library is used: ln from `decimal`, pi by Machin's
formula, sin and cos by their Taylor series at 60 digits."""
from decimal import Decimal, getcontext

getcontext().prec = 60


def atan_inv(n):
    """arctan(1/n) by its Taylor series."""
    x = Decimal(1) / n
    x2 = x * x
    s, t, k = Decimal(0), x, 0
    while True:
        term = t / (2 * k + 1)
        if abs(term) < Decimal(10) ** -58:
            break
        s += -term if k % 2 else term
        t *= x2
        k += 1
    return s


def pi():
    return 16 * atan_inv(5) - 4 * atan_inv(239)  # Machin


def sincos(x):
    s, c = Decimal(0), Decimal(0)
    t, k = Decimal(1), 0  # x^k / k!
    while True:
        if k > 0:
            t = t * x / k
        if k > 4 and abs(t) < Decimal(10) ** -58:
            break
        r = k % 4
        if r == 0:
            c += t
        elif r == 1:
            s += t
        elif r == 2:
            c -= t
        else:
            s -= t
        k += 1
    return s, c


def split(v, n=2):
    out = []
    for _ in range(n):
        h = float(v)  # round to nearest double
        out.append(h)
        v -= Decimal(h)
    return out


def hx(v):
    return float.hex(v)


def main():
    P = pi()
    print("// Generated by gen_gauss_tables.py (60-digit decimal evaluation, split into doubles")
    print("// rounded to nearest); the constants of gauss_cr.hpp. Do not edit.")
    print("#pragma once")
    print("namespace ghx {")
    print("namespace gcr {")
    l2 = split(Decimal(2).ln(), 3)
    print(f"constexpr double kLn2[3] = {{{', '.join(hx(v) for v in l2)}}};")
    p2 = split(P / 2, 3)
    print(f"constexpr double kPio2[3] = {{{', '.join(hx(v) for v in p2)}}};")
    print(f"constexpr double kTwoOverPi = {hx(float(2 / P))};")
    for name, v in (("kThird", Decimal(1) / 3), ("kFifth", Decimal(1) / 5), ("kSixth", Decimal(1) / 6),
                    ("kInv24", Decimal(1) / 24), ("kInv120", Decimal(1) / 120)):
        a, b = split(v)
        print(f"constexpr double {name}[2] = {{{hx(a)}, {hx(b)}}};")
    print("// log(j/128), j = 96..192: {hi, lo}")
    print("constexpr double kLogC[97][2] = {")
    for j in range(96, 193):
        a, b = split((Decimal(j) / 128).ln())
        print(f"    {{{hx(a)}, {hx(b)}}},")
    print("};")
    print("// sin(j/64) and cos(j/64), j = 0..51: {sin hi, sin lo, cos hi, cos lo}")
    print("constexpr double kSinCos[52][4] = {")
    for j in range(52):
        s, c = sincos(Decimal(j) / 64)
        sa, sb = split(s)
        ca, cb = split(c)
        print(f"    {{{hx(sa)}, {hx(sb)}, {hx(ca)}, {hx(cb)}}},")
    print("};")
    print("}  // namespace gcr")
    print("}  // namespace ghx")


if __name__ == "__main__":
    main()
