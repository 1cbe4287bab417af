"""The quadratic squeeze of the ZIGNOR wedge test (csrc/gen_zig_tables.py, FOHIP_ZIG_SQUEEZE_INIT),
pinned against 40-digit decimal arithmetic and against the exact binary64 wedge test.

For a draw rejected by layer i's fast test (|value| in [X[i+1], X[i])) the kernels decide the
wedge test F[i+1] + (F[i] - F[i+1]) u < exp(-value^2 / 2) (random/normal.rs:74-77) through
q(t) = t + c t (t - 1), t = (|value| - X[i+1]) / w: u - q > gu accepts, u - q < -gl rejects, and only
the band between runs the exponential. That is exact only if the curve psi(t) lies inside
[q - gl, q + gu] with room for the binary32 evaluation, which is what these tests check."""

from __future__ import annotations

import math
import re
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
INC = ROOT / "finoptions-engine_amd" / "csrc" / "zig_tables.inc"


def _block(name):
    text = INC.read_text()
    body = text[text.index(f"#define {name}"):]
    body = body[body.index("{") + 1:body.index("\n}")]
    return body


def _tables():
    xs = [float.fromhex(v) for v in re.findall(r"(-?0x[0-9a-fp.+-]+)", _block("FOHIP_ZIG_X_INIT"))]
    fs = [float.fromhex(v) for v in re.findall(r"(-?0x[0-9a-fp.+-]+)", _block("FOHIP_ZIG_F_INIT"))]
    sq = []
    for row in re.findall(r"\{([^}]*)\}", _block("FOHIP_ZIG_SQUEEZE_INIT")):
        vals = [v.strip().rstrip("f") for v in row.split(",")]
        sq.append(tuple(0.0 if v in ("0.0", "0") else float.fromhex(v) for v in vals))
    return xs, fs, sq


# What the squeeze must leave beyond the curve: the kernel's binary32 arithmetic deviates from
# exact q(t) and u by less than 1e-6, the exact test's binary64 rounding by less than 2e-13.
SLACK = 2e-6


def test_squeeze_bounds_hold_in_decimal_arithmetic():
    getcontext().prec = 40
    xs, fs, sq = _tables()
    assert len(xs) == 257 and len(fs) == 257 and len(sq) == 256
    for i in range(1, 256):
        inv_w, c, gu, gl = sq[i]
        a, b = Decimal(xs[i + 1]), Decimal(xs[i])
        w, h = b - a, Decimal(fs[i + 1]) - Decimal(fs[i])
        assert abs(float(1 / w) - inv_w) <= 1e-7 * inv_w
        for k in range(0, 401):
            t = Decimal(k) / 400
            x = a + t * w
            psi = (Decimal(fs[i + 1]) - (-(x * x) / 2).exp()) / h
            q = t + Decimal(c) * t * (t - 1)
            assert float(psi - q) <= gu - SLACK, (i, k)
            assert float(q - psi) <= gl - SLACK, (i, k)


def test_squeeze_decisions_agree_with_the_exact_wedge_test():
    """The kernel's binary32 squeeze, emulated with numpy, against the exact binary64 test on
    random rejected draws of every layer; every decided draw must agree."""
    xs, fs, sq = _tables()
    rng = np.random.default_rng(7)
    decided = undecided = 0
    for i in range(1, 256):
        inv_w, c, gu, gl = (np.float32(v) for v in sq[i])
        a, b = xs[i + 1], xs[i]
        # |value| uniform over the layer's wedge, u over its 53-bit grid
        x = a + (b - a) * rng.random(4000)
        x = x[(x >= a) & (x < b)]
        ubits = rng.integers(0, 1 << 63, size=x.size, dtype=np.int64).astype(np.uint64) << np.uint64(1)
        u = (ubits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        t = (x - a).astype(np.float32) * inv_w
        q = np.float32(c) * (t * (t - np.float32(1.0))) + t   # an FMA in the kernel; the
        # separate rounding here is within the slack
        uf = (ubits >> np.uint64(41)).astype(np.uint32).astype(np.float32) * np.float32(2.0 ** -23)
        d = uf - q
        ordinate = fs[i + 1] + (fs[i] - fs[i + 1]) * u
        exact = ordinate < np.exp(-0.5 * x * x)
        acc, rej = d > gu, d < -gl
        assert not np.any(acc & ~exact), i
        assert not np.any(rej & exact), i
        decided += int(np.sum(acc | rej))
        undecided += int(np.sum(~(acc | rej)))
    assert undecided < 0.01 * (decided + undecided)


def test_band_is_narrow_for_the_normal_stream():
    """The exponential runs only inside the band: the probability per normal, from the layers'
    fast-test failure rates and band widths, is about 1e-5."""
    xs, _, sq = _tables()
    p = sum((1.0 - xs[i + 1] / xs[i]) / 256.0 * (sq[i][2] + sq[i][3]) for i in range(1, 256))
    assert p < 3e-5
