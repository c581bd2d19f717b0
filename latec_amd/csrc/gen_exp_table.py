"""Generate exp_table.h: the 2^(i/128) table of the table-driven exp.

Entry i holds T_i = RN(2^(i/128)) (stored minus i<<45 so that adding the
scaled integer part k<<45 rebuilds the exponent) and tail_i =
RN((2^(i/128) - T_i) / T_i), computed here at 80 decimal digits.  This is the
table layout of the table-driven exp of ARM's optimized-routines (MIT), which
glibc ships as its double exp; latec_exp() in ddmath.h evaluates that
algorithm so that the GPU-built rANS tables match the host libm bit for bit.

    python latec_amd/csrc/gen_exp_table.py > latec_amd/csrc/exp_table.h
"""

import struct
from decimal import Decimal, getcontext

getcontext().prec = 80
N = 128


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def main():
    rows = []
    two = Decimal(2)
    for i in range(N):
        exact = two ** (Decimal(i) / Decimal(N))
        t = float(exact)  # correctly rounded (Decimal -> float is RN)
        tail = float((exact - Decimal(t)) / Decimal(t))
        rows.append((bits(tail), (bits(t) - (i << 45)) & 0xFFFFFFFFFFFFFFFF))
    print("// Generated by gen_exp_table.py -- do not edit.")
    print("#pragma once")
    print("#include <stdint.h>")
    print("#define LATEC_EXP_TABLE_BITS 7")
    print("static const uint64_t latec_exp_tab[2 * 128] = {")
    for a, b in rows:
        print(f"  0x{a:016x}ull, 0x{b:016x}ull,")
    print("};")


if __name__ == "__main__":
    main()
