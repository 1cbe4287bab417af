"""Generate zig_tables.inc: the 256-layer ZIGNOR normal tables as exact hex-float literals.

Construction (Doornik 2005, "An Improved Ziggurat Method to Generate Normal Random Samples"):
f(x) = exp(-x^2/2), tail start r = 3.6541528853610088, common layer area v = 0.00492867323399;
x_0 = v / f(r), x_1 = r, x_i = f^-1(v / x_{i-1} + f(x_{i-1})) for i = 2..255, x_256 = 0;
F_i = f(x_i). Each value is rounded to 18 decimal places, the rendering in which the tables were
published for the Rust `rand` family (and used by the reference, random/normal_tables.rs).
Entry F_94 is taken as published (0.206405406398679298): the publishing platform's exp() put it
one unit lower in the 18th decimal than glibc's does.
"""

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
    out.write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
