"""Generate zig_tables.inc: the 256-layer ZIGNOR normal tables as exact hex-float literals.

Construction (Doornik 2005, "An Improved Ziggurat Method to Generate Normal Random Samples"):
f(x) = exp(-x^2/2), tail start r = 3.6541528853610088, common layer area v = 0.00492867323399;
x_0 = v / f(r), x_1 = r, x_i = f^-1(v / x_{i-1} + f(x_{i-1})) for i = 2..255, x_256 = 0;
F_i = f(x_i). Each value is rounded to 18 decimal places, the rendering in which the tables were
published for the Rust `rand` family (and used by the reference, random/normal_tables.rs).
Entry F_94 is taken as published (0.206405406398679298): the publishing platform's exp() put it
one unit lower in the 18th decimal than glibc's does.

It also writes the wedge-test squeeze (FOHIP_ZIG_SQUEEZE_INIT, one binary32 quadruple per layer),
see squeeze() below.
"""

import struct
from math import exp, log, sqrt
from pathlib import Path

R = 3.6541528853610088
V = 0.00492867323399


def tables() -> tuple[list[float], list[float], float]:
    f = lambda x: exp(-x * x / 2.0)
    x = [0.0] * 257
    x[0] = V / f(R)
    x[1] = R
    for i in range(2, 256):
        x[i] = sqrt(-2.0 * log(V / x[i - 1] + f(x[i - 1])))
    x[256] = 0.0
    dec = lambda v: float("%.18f" % v)
    xs = [dec(v) for v in x]
    fs = [dec(f(v)) for v in x]
    fs[94] = float("0.206405406398679298")
    return xs, fs, dec(R)


def _f32(v: float) -> float:
    return struct.unpack("<f", struct.pack("<f", v))[0]


def _f32_up(v: float) -> float:
    """The smallest binary32 value >= v (v >= 0)."""
    r = _f32(v)
    if r < v:
        r = struct.unpack("<f", struct.pack("<I", struct.unpack("<I", struct.pack("<f", r))[0] + 1))[0]
    return r


# Margin of the squeeze in units of the layer's ordinate range: covers the binary32 evaluation of
# t, q and u in the kernel (below 1e-6 together), the binary64 rounding of the exact test's
# ordinate and exponential (below 2e-13 in these units), and the sampling of the residual below.
SQUEEZE_MARGIN = 1e-5
SQUEEZE_SAMPLES = 2000


def squeeze(xs: list[float], fs: list[float]) -> list[tuple[float, float, float, float]]:
    """Quadratic squeeze of the ZIGNOR wedge test, one (1/w, c, gu, gl) per layer 1..255.

    A draw rejected by the fast test of layer i has x = |value| in [a, b) = [x_{i+1}, x_i) and
    ordinate F_{i+1} - h u with h = F_{i+1} - F_i and u the standard uniform of the next draw; it is
    accepted iff the ordinate lies below exp(-x^2/2). With t = (x - a) / w, w = b - a, the curve is
    F_{i+1} - h psi(t), so the test reads u > psi(t). On [0, 1] psi stays within [q - gl, q + gu] of
    q(t) = t + c t (t - 1), the least-squares quadratic through psi's end points (c rounded to
    binary32; gu and gl are the largest deviations on a 2000-interval grid plus SQUEEZE_MARGIN,
    rounded up to binary32). u - q(t) > gu therefore accepts and u - q(t) < -gl rejects exactly
    as the exponential would; only the narrow band between needs the exponential (about 1e-5 of
    all normals). Layer 0 is the tail and is not squeezed."""
    out = [(0.0, 0.0, 0.0, 0.0)]
    for i in range(1, 256):
        a, b = xs[i + 1], xs[i]
        w, h = b - a, fs[i + 1] - fs[i]
        ts = [k / SQUEEZE_SAMPLES for k in range(SQUEEZE_SAMPLES + 1)]
        psi = [(fs[i + 1] - exp(-(a + t * w) * (a + t * w) / 2.0)) / h for t in ts]
        num = sum((p - t) * t * (t - 1.0) for t, p in zip(ts, psi))
        den = sum((t * (t - 1.0)) ** 2 for t in ts)
        c = _f32(num / den)
        dev = [p - (t + c * t * (t - 1.0)) for t, p in zip(ts, psi)]
        gu = _f32_up(max(0.0, max(dev)) + SQUEEZE_MARGIN)
        gl = _f32_up(max(0.0, -min(dev)) + SQUEEZE_MARGIN)
        out.append((_f32(1.0 / w), c, gu, gl))
    return out


def _c_float(v: float) -> str:
    return "0.0f" if v == 0.0 else f"{v.hex()}f"


def main() -> None:
    xs, fs, r = tables()
    out = Path(__file__).with_name("zig_tables.inc")
    lines = ["// Generated by gen_zig_tables.py (ZIGNOR, 256 layers). Do not edit.",
             f"#define FOHIP_ZIG_R {r.hex()}"]
    lines.append("#define FOHIP_ZIG_X_INIT { \\")
    lines += [f"    {v.hex()}, \\" for v in xs]
    lines.append("}")
    lines.append("#define FOHIP_ZIG_F_INIT { \\")
    lines += [f"    {v.hex()}, \\" for v in fs]
    lines.append("}")
    lines.append("#define FOHIP_ZIG_SQUEEZE_INIT { \\")
    lines += ["    {" + ", ".join(_c_float(v) for v in q) + "}, \\" for q in squeeze(xs, fs)]
    lines.append("}")
    out.write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
