"""Generates clm_math_tables.h: the tables of the FP64 log/exp/pow in clm_math.h.

Run from this directory:  python3 gen_clm_math.py > clm_math_tables.h

The values are computed with Python's decimal module at 80 significant digits
and rounded once to the nearest double (Decimal -> float conversion is
correctly rounded), so the header is reproducible bit for bit.

Logarithm (Tang 1990, table-driven, 256 subintervals). For x = 2^k * z with
z in [OFF, 2*OFF), OFF ~ 1/sqrt(2), the top 8 bits of (bits(z) - LOG_OFF) pick
subinterval i. Each entry holds
  invc   1/c for a point c of the subinterval with 9 significant bits (the one
         of the two nearest to 1/midpoint that gives the smaller |r|), so that
         r = fma(z, invc, -1) is exact: z has 53 significant bits, z*invc is a
         multiple of 2^-61 and |r| < 2^-8, so r fits in 53 bits (checked below);
  logc   -log(invc) = log(c) as a double-double (hi, lo).
The subinterval that contains 1.0 has invc = 1 exactly (log(c) = 0), so for x
near 1 the result is log1p(r) with r = x - 1 exact.

Exponential (Tang 1989, 128 subintervals): 2^(j/128) = hi + lo for j = 0..127.
"""
from __future__ import annotations

import struct
from decimal import Decimal, getcontext
from fractions import Fraction

getcontext().prec = 80

N_LOG = 256
N_EXP = 128


def bits(d: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def from_bits(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def dec(q: Fraction) -> Decimal:
    return Decimal(q.numerator) / Decimal(q.denominator)


def nearest(q) -> float:
    """Nearest double of a Fraction or Decimal (ties to even)."""
    if isinstance(q, Fraction):
        q = dec(q)
    return float(q)


def split(q: Decimal) -> tuple[float, float]:
    hi = nearest(q)
    lo = nearest(q - Decimal(hi))
    return hi, lo


def hexf(x: float) -> str:
    if x == 0.0:
        return "0x0p+0"
    return float.hex(x)


def sig_bits(q: Fraction) -> int:
    """Number of significant bits of a dyadic rational."""
    n, d = q.numerator, q.denominator
    assert d & (d - 1) == 0, "not dyadic"
    n = abs(n)
    while n % 2 == 0:
        n //= 2
    return n.bit_length()


LN2 = Decimal(2).ln()


def log_table():
    # LOG_OFF: close to bits(1/sqrt(2)), with bits 0..43 equal to 2^43 so that
    # 1.0 = bits 0x3ff0000000000000 lies in the middle of its subinterval.
    base = bits(0.7071067811865476)
    off = (base & ~((1 << 44) - 1)) | (1 << 43)
    rows = []
    center = None
    for i in range(N_LOG):
        lo_b = off + (i << 44)
        hi_b = off + ((i + 1) << 44)
        z_lo = Fraction(from_bits(lo_b))
        z_hi = Fraction(from_bits(hi_b))   # exclusive
        if z_lo <= 1 < z_hi:
            invc = Fraction(1)
            center = i
        else:
            inv = 1 / ((z_lo + z_hi) / 2)
            quantum = Fraction(1, 256) if inv >= 1 else Fraction(1, 512)
            cands = (Fraction(int(inv / quantum)) * quantum,
                     Fraction(int(inv / quantum) + 1) * quantum)
            invc = min(cands, key=lambda c: max(abs(z_lo * c - 1), abs(z_hi * c - 1)))
        assert sig_bits(invc) <= 9
        # |z * invc - 1| < 2^-8 on the whole subinterval (linear in z)
        rmax = max(abs(z_lo * invc - 1), abs(z_hi * invc - 1))
        assert rmax < Fraction(1, 256), (i, float(rmax))
        logc = -(dec(invc).ln())
        hi, lo = split(logc)
        rows.append((nearest(invc), hi, lo, float(rmax)))
    return off, rows, center


def exp_table():
    rows = []
    for j in range(N_EXP):
        t = (Decimal(j) / Decimal(N_EXP) * LN2).exp()
        rows.append(split(t))
    return rows


def round_bits(q: Decimal, nbits: int) -> float:
    """q rounded to nbits significant bits (as a double)."""
    f = Fraction(q)
    e = 0
    a = abs(f)
    while a >= 2:
        a /= 2
        e += 1
    while a < 1:
        a *= 2
        e -= 1
    quantum = Fraction(2) ** (e - nbits + 1)
    return float(Fraction(round(f / quantum)) * quantum)


def main():
    off, lrows, center = log_table()
    erows = exp_table()
    ln2_hi = round_bits(LN2, 42)          # k * ln2_hi exact for |k| < 2^11
    ln2_lo = nearest(LN2 - Decimal(ln2_hi))
    ln2n = LN2 / N_EXP
    ln2n_hi = round_bits(ln2n, 35)        # kd * ln2n_hi exact for |kd| < 2^18
    ln2n_lo = nearest(ln2n - Decimal(ln2n_hi))
    invln2n = nearest(Decimal(N_EXP) / LN2)
    inv_ln10 = Decimal(1) / Decimal(10).ln()
    il10_hi, il10_lo = split(inv_ln10)
    out = []
    w = out.append
    w("/* clm_math_tables.h — generated by gen_clm_math.py; do not edit. */")
    w("#ifndef CLM_MATH_TABLES_H")
    w("#define CLM_MATH_TABLES_H")
    w(f"#define CLM_LOG_OFF 0x{off:016x}ULL")
    w(f"#define CLM_LOG_CENTER {center}")
    w(f"#define CLM_LN2_HI {hexf(ln2_hi)}")
    w(f"#define CLM_LN2_LO {hexf(ln2_lo)}")
    w(f"#define CLM_LN2N_HI {hexf(ln2n_hi)}")
    w(f"#define CLM_LN2N_LO {hexf(ln2n_lo)}")
    w(f"#define CLM_INVLN2N {hexf(invln2n)}")
    w(f"#define CLM_INVLN10_HI {hexf(il10_hi)}")
    w(f"#define CLM_INVLN10_LO {hexf(il10_lo)}")
    w(f"/* max |z*invc - 1| over all subintervals: {max(r[3] for r in lrows):.6g} (< 2^-8) */")
    w("/* {invc, log(c) hi, log(c) lo} for the 256 subintervals of [OFF, 2*OFF) */")
    w("CLM_MATH_TAB double clm_log_tab[256][3] = {")
    for invc, hi, lo, _ in lrows:
        w(f"    {{{hexf(invc)}, {hexf(hi)}, {hexf(lo)}}},")
    w("};")
    w("/* 2^(j/128) = hi + lo */")
    w("CLM_MATH_TAB double clm_exp_tab[128][2] = {")
    for hi, lo in erows:
        w(f"    {{{hexf(hi)}, {hexf(lo)}}},")
    w("};")
    w("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()
