"""The constants of the GBM step's table-driven exponential, pinned against 60-digit decimal
arithmetic: 2^(j/32) as a leading (correctly rounded) and trailing binary64 pair whose sum carries
about 106 significant bits, and ln2/32 split so that n * L1 is exact for |n| < 2^20."""

from __future__ import annotations

import re
import struct
from decimal import Decimal, getcontext
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
INC = ROOT / "finoptions-engine_amd" / "csrc" / "exp_table.inc"


def _parse():
    text = INC.read_text()
    scalars = {k: float.fromhex(v) for k, v in re.findall(r"#define FOHIP_EXP_(\w+) (\S+)", text)
               if k != "TABLE_INIT"}
    pairs = [(float.fromhex(a), float.fromhex(b))
             for a, b in re.findall(r"\{(0x[^,]+), (-?0x[^}]+)\}", text)]
    return scalars, pairs


def test_table_entries_are_two_part_powers_of_two():
    getcontext().prec = 60
    _, pairs = _parse()
    assert len(pairs) == 32
    for j, (hi, lo) in enumerate(pairs):
        exact = Decimal(2) ** (Decimal(j) / 32)
        assert hi == float(exact)                       # leading part correctly rounded
        assert abs(lo) <= abs(hi) * 2.0 ** -53          # trailing part below half an ulp of it
        err = abs(Decimal(hi) + Decimal(lo) - exact) / exact
        assert err < Decimal(2) ** -104, (j, err)


def test_reduction_constants_split_ln2_over_32():
    getcontext().prec = 60
    sc, _ = _parse()
    step = Decimal(2).ln() / 32
    l1, l2 = sc["L1"], sc["L2"]
    bits = struct.unpack("<Q", struct.pack("<d", l1))[0]
    assert bits & ((1 << 20) - 1) == 0                  # n * L1 exact for |n| < 2^20
    # L2 is the binary64 rounding of the remainder: the split is exact to half an ulp of L2
    assert abs(Decimal(l1) + Decimal(l2) - step) <= abs(Decimal(l2)) * Decimal(2) ** -53
    assert sc["INV_L"] == float(Decimal(32) / Decimal(2).ln())
