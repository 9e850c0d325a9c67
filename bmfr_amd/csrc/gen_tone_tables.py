"""Writes tone_tables.h, the tables and coefficients of tone_curve.h, for
the power v^y with y = (double)0.454545f (the float nearest 0.454545):

  kToneLog[i], i = 0..255, for m in [0.5 + i/512, 0.5 + (i+1)/512):
      c_i  = 1 / m_mid rounded to 12 significant bits (so that m * c_i - 1
             is exact in FP64 for every m of at most 24 bits), and
      c_i^-y, correctly rounded to double;
  kTonePow2[k], k = 0..148:  2^(-y k), correctly rounded to double;
  kToneBinom2..5: the binomial coefficients C(y, k) = y (y-1) ... (y-k+1) / k!,
      correctly rounded to double.

Values are computed with the decimal module at 60 significant digits and
converted with float(), which rounds correctly; they are written as
hexadecimal floating literals, so the header holds the exact doubles.

    python3 gen_tone_tables.py > tone_tables.h
"""
import struct
from decimal import Decimal, getcontext

getcontext().prec = 60
# the exponent of the reference's tone curve, 0.454545f, as the float it is
Y = Decimal(struct.unpack("<f", struct.pack("<f", 0.454545))[0])
LN2 = Decimal(2).ln()


def log_entry(i):
    m_mid = Decimal(1) / 2 + (Decimal(i) + Decimal(1) / 2) / 512
    j = int((Decimal(2048) / m_mid).to_integral_value())   # c = j / 2^11, 12 significant bits
    assert 2048 <= j < 4096
    c = Decimal(j) / 2048
    return float(c), float((-Y * c.ln()).exp())


def pow2_entry(k):
    return float((-Y * k * LN2).exp())


def binom(k):
    num = Decimal(1)
    for t in range(k):
        num *= Y - t
    den = 1
    for t in range(2, k + 1):
        den *= t
    return float(num / den)


def main():
    out = ["// Generated by gen_tone_tables.py: the tables of tone_curve.h.",
           "#pragma once", "",
           "// {c_i, c_i^-y}: c_i = 1 / (0.5 + (i + 0.5) / 512) to 12 significant bits",
           "BMFR_TONE_TABLE const double kToneLog[256][2] BMFR_TONE_ALIGN = {"]
    for i in range(256):
        c, cy = log_entry(i)
        out.append(f"    {{{c.hex()}, {cy.hex()}}},")
    out += ["};", "", "// 2^(-y k), k = 0..148",
            "BMFR_TONE_TABLE const double kTonePow2[149] BMFR_TONE_ALIGN = {"]
    for k in range(149):
        out.append(f"    {pow2_entry(k).hex()},")
    out += ["};", "", "// binomial coefficients C(y, k)"]
    for k in range(2, 6):
        out.append(f"constexpr double kToneBinom{k} = {binom(k).hex()};")
    out.append("")
    print("\n".join(out))


if __name__ == "__main__":
    main()
