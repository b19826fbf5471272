"""Writes log_f32_table.h, the reduction table of LogOfFloat (log_f32.h).

Entry i (0 <= i < 128) covers mantissas z in [1 + i/128, 1 + (i+1)/128):
  invc: 1 / c_i rounded to 24 significant bits (a float), c_i the interval's
        centre, so that z * invc - 1 is exact for a 24-bit z; invc = 1 for
        i = 0 (z near 1 from above) and 0.5 for i = 127 (z near 2: the
        argument just below a power of two, whose exponent LogOfFloat raises
        by one instead of adding log 2 here, so that below 1 nothing cancels);
  hi, lo: -log(invc) as a double plus a float correction, to ~2^-77 relative
        (0 for i = 0 and i = 127).

Run: python make_log_table.py > log_f32_table.h
"""

from __future__ import annotations

import struct

import mpmath

mpmath.mp.prec = 200

# Head of ln 2 with 32 significant bits: k * LN2_HI is exact for |k| < 2^21.
LN2 = mpmath.log(2)


def to_float(x) -> float:
    return struct.unpack("<f", struct.pack("<f", float(x)))[0]


def head_bits(x: float, bits: int) -> float:
    m, e = mpmath.frexp(mpmath.mpf(x))
    scale = mpmath.mpf(2) ** bits
    return float(mpmath.nint(m * scale) / scale * mpmath.mpf(2) ** e)


LN2_HI = head_bits(float(LN2), 32)
LN2_LO = float(LN2 - LN2_HI)


def entry(i: int):
    if i == 0:
        return 1.0, 0.0, 0.0
    if i == 127:
        return 0.5, 0.0, 0.0
    c = 1 + (mpmath.mpf(i) + mpmath.mpf("0.5")) / 128
    invc = to_float(1 / c)
    exact = -mpmath.log(mpmath.mpf(invc))
    hi = float(exact)
    lo = to_float(exact - hi)
    return invc, hi, lo


def main() -> None:
    print("// Generated by make_log_table.py: reduction table of LogOfFloat.")
    print("// {-log(invc) head (double), correction (float), invc (float)}")
    print("#define THMM_LOG_F32_TABLE \\")
    print("  { \\")
    for i in range(128):
        invc, hi, lo = entry(i)
        sep = "," if i < 127 else ""
        print(f"    {{{float(hi).hex()}, {float(lo).hex()}f, {float(invc).hex()}f}}{sep} \\")
    print("  }")
    print(f"#define THMM_LN2_HI {LN2_HI.hex()}")
    print(f"#define THMM_LN2_LO {LN2_LO.hex()}")


if __name__ == "__main__":
    main()
