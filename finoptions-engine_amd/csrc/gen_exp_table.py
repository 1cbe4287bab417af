"""Generate exp_table.inc: constants of the table-driven binary64 exponential of the GBM step.

Method: P. T. P. Tang, "Table-driven implementation of the exponential function in IEEE
floating-point arithmetic", ACM TOMS 15(2), 1989, with 32 table entries. x = (32 m + j) ln2/32 + r,
|r| <= ln2/64, exp(x) = 2^m * 2^(j/32) * exp(r). The reduction is Cody-Waite: ln2/32 split into a
leading part with its low 20 significand bits cleared (so n * L1 is exact for |n| < 2^20) and the
binary64 rounding of the remainder. 2^(j/32) is stored as a leading binary64 value (its correct
rounding) and a trailing one (the rounding of the remainder). Values are computed with 60-digit
decimal arithmetic and written as exact hex-float literals.
"""

from __future__ import annotations

import struct
from decimal import Decimal, getcontext
from pathlib import Path

getcontext().prec = 60


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _from_bits(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def constants() -> dict:
    ln2 = Decimal(2).ln()
    step = ln2 / 32
    l1 = _from_bits(_bits(float(step)) & ~((1 << 20) - 1))
    l2 = float(step - Decimal(l1))
    inv = float(Decimal(32) / ln2)
    table = []
    for j in range(32):
        v = Decimal(2) ** (Decimal(j) / 32)
        hi = float(v)
        table.append((hi, float(v - Decimal(hi))))
    return {"l1": l1, "l2": l2, "inv": inv, "table": table}


def main() -> None:
    c = constants()
    lines = ["// Generated by gen_exp_table.py (Tang's table-driven exp, 32 entries). Do not edit.",
             f"#define FOHIP_EXP_INV_L {c['inv'].hex()}",
             f"#define FOHIP_EXP_L1 {c['l1'].hex()}",
             f"#define FOHIP_EXP_L2 {c['l2'].hex()}",
             "#define FOHIP_EXP_TABLE_INIT { \\"]
    lines += [f"    {{{hi.hex()}, {lo.hex()}}}, \\" for hi, lo in c["table"]]
    lines += ["}", ""]
    out = Path(__file__).with_name("exp_table.inc")
    text = "\n".join(lines)
    if not out.exists() or out.read_text() != text:
        out.write_text(text)


if __name__ == "__main__":
    main()
