"""Write tests/golden/reference_vectors.json from the reference's own test vectors.

Reads the reference sources as text (no build, no import) and extracts the known-answer data the
reference's tests assert on:
  * Philox4x32-10 known-answer vectors       crates/finoptions-sim/src/random/philox.rs:121-130
  * 16-value ZIGNOR bit pattern              crates/finoptions-sim/src/random/normal.rs:116-141
  * ziggurat table digest (X, F, R bits)     crates/finoptions-sim/src/random/normal_tables.rs
  * HHL one-dividend reference value         crates/finoptions-lsm/src/policy.rs:1617-1670
Run once in a container that mounts the reference at /root/reference; the JSON is committed.
"""

from __future__ import annotations

import hashlib
import json
import re
import struct
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).with_name("reference_vectors.json")


def hexes(text: str) -> list[int]:
    return [int(h.replace("_", ""), 16) for h in re.findall(r"0x([0-9a-f_]+)", text)]


def main() -> None:
    philox = (REF / "crates/finoptions-sim/src/random/philox.rs").read_text()
    block = philox[philox.index("fn matches_random123_known_answer_vectors"):]
    block = block[: block.index("#[test]", 10)]
    words = hexes(block)
    # two vectors: counter, key, expected (counter/key of the second use u32::MAX literals)
    kats = [
        {"counter": [0, 0, 0, 0], "key": [0, 0], "expected": words[0:4]},
        {"counter": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2, "expected": words[4:8]},
    ]
    normal = (REF / "crates/finoptions-sim/src/random/normal.rs").read_text()
    nb = normal[normal.index("fn pins_the_standard_normal_transform_sequence"):]
    nb = nb[: nb.index("#[test]", 10)]
    nh = hexes(nb)
    seed, domain = nh[0], nh[1]
    path = int(re.search(r"PhiloxPathStream::new\([^,]+,\s*[^,]+,\s*(\d+),", nb).group(1))
    count = int(re.search(r"PhiloxPathStream::new\([^,]+,\s*[^,]+,\s*\d+,\s*(\d+)\)", nb).group(1))
    normals = nh[2:2 + count]
    tables = (REF / "crates/finoptions-sim/src/random/normal_tables.rs").read_text()

    def table(name: str) -> list[float]:
        m = re.search(name + r": \[f64; 257\] =\s*\[(.*?)\];", tables, re.S)
        return [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]

    x, f = table("ZIG_NORM_X"), table("ZIG_NORM_F")
    r = float(re.search(r"ZIG_NORM_R: f64 = ([0-9.]+);", tables).group(1))
    digest = hashlib.sha256(struct.pack("<257d", *x) + struct.pack("<257d", *f)
                            + struct.pack("<d", r)).hexdigest()
    policy = (REF / "crates/finoptions-lsm/src/policy.rs").read_text()
    hhl = float(re.search(r"const REFERENCE: f64 = ([0-9_.]+);", policy).group(1).replace("_", ""))
    t99 = float(re.search(r"const STUDENT_T_99_DF_11: f64 = ([0-9_.]+);", policy).group(1).replace("_", ""))
    OUT.write_text(json.dumps({
        "philox_kats": kats,
        "zignor_sequence": {"seed": seed, "domain": domain, "path": path,
                            "bits": [f"0x{v:016x}" for v in normals]},
        "zig_table_sha256": digest,
        "zig_f94": f[94],
        "hhl_one_dividend_reference": hhl,
        "student_t_99_df_11": t99,
    }, indent=2) + "\n")
    print(OUT)


if __name__ == "__main__":
    main()
