"""Transcribe the reference's known-answer vectors into JSON fixtures.

Reads, as text, the constant arrays of the reference's own unit tests:
  core/src/common/mont/tests/main.cu    (BN254 Fr big-int / Montgomery KATs,
                                          instance1 :237-253, instance2 :370-385,
                                          to/from Montgomery :500-509,
                                          mm2 ops instance3 :527-549)
  core/src/common/curve/tests/bn254.cu  (G1 points in Montgomery form :177-203)
and writes field_kat.json / curve_kat.json next to this script.  Values are
stored as big-endian hex strings of the integer the limb array encodes
(BIG_INTEGER_CHUNKS8 lists the most significant 32-bit chunk first).

Run once in a container that has the reference at /root/reference:
    python tests/golden/make_kat_fixtures.py
The JSON files are committed; nothing at test time reads the reference.
"""
import json
import os
import re

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

ARR = re.compile(
    r"const u32 (\w+)\[[^\]]*\]\s*=\s*BIG_INTEGER_CHUNKS(8|16)\(([^)]*)\);")


def chunks_to_hex(body):
    parts = [p.strip() for p in body.split(",")]
    value = 0
    for p in parts:
        value = (value << 32) | int(p, 16)
    return hex(value)


def parse(path, start_marker=None):
    text = open(path).read()
    blocks = {}
    current = "top"
    for line in text.splitlines():
        m = re.match(r"\s*namespace (instance\d+)", line)
        if m:
            current = m.group(1)
        if "TEST_CASE(\"Convert to and from Montgomery\")" in line:
            current = "convert"
        for mm in ARR.finditer(line):
            blocks.setdefault(current, {})[mm.group(1)] = chunks_to_hex(mm.group(3))
    return blocks


def main():
    mont = parse(os.path.join(REF, "core/src/common/mont/tests/main.cu"))
    curve = parse(os.path.join(REF, "core/src/common/curve/tests/bn254.cu"))
    field = {
        "source": "core/src/common/mont/tests/main.cu (bn254_fr::Element)",
        "semantics": {
            "sum": "a + b as 256-bit integers (bn_add)",
            "sum_mont": "Element a + Element b (mod r)",
            "sub": "a - b mod 2^256 (bn_sub)",
            "sub_mont": "Element a - Element b (mod r)",
            "prod": "a * b as 512-bit integers",
            "prod_mont": "Element(a) * Element(b) = a*b*R^-1 mod r",
            "a_square": "a^2 as 512-bit integer",
            "a_square_mont": "from_number(a) squared, Montgomery",
            "pow_mont": "from_number(a).pow(b), Montgomery",
            "a_inv_mont": "from_number(a).invert(), Montgomery",
            "a_slr125": "a >> 125",
            "x_mont": "from_number(x)",
            "r_sub": "x - y modulo 2r (sub_modulo_mm2)",
            "r_add": "x + y modulo 2r (add_modulo_mm2)",
            "x_mod_m": "x reduced once modulo r (modulo_m)",
        },
        "instances": mont,
    }
    pts = curve.get("top", {})
    curve_out = {
        "source": "core/src/common/curve/tests/bn254.cu:177-203 (Montgomery form, Fq)",
        "relations": {
            "pr": "p1 + p2",
            "pr100": "p1 + 100 * p2",
            "p1r128": "128 * p1",
        },
        "points": pts,
    }
    with open(os.path.join(HERE, "field_kat.json"), "w") as f:
        json.dump(field, f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "curve_kat.json"), "w") as f:
        json.dump(curve_out, f, indent=1, sort_keys=True)
    print("instances:", {k: sorted(v) for k, v in mont.items()})
    print("points:", sorted(pts))


if __name__ == "__main__":
    main()
