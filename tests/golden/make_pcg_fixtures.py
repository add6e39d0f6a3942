"""Writes the PCG32 fixtures with a pure-Python PCG32, independent of the C
oracle, so the oracle test compares two implementations.

pcg32_kat.json holds the published pcg32-demo vector of the PCG minimal C
reference (srandom(42, 54), the first block the demo prints); this script re-derives it and refuses to
write anything if it disagrees.
"""
import hashlib
import json
import struct
from pathlib import Path

M = 6364136223846793005
MASK = (1 << 64) - 1
HERE = Path(__file__).resolve().parent
PUBLISHED = ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b", "0xcbed606e"]


def pcg32(n, seed, seq):
    inc = ((seq << 1) | 1) & MASK
    s = (0 * M + inc) & MASK
    s = (s + seed) & MASK
    s = (s * M + inc) & MASK
    out = []
    for _ in range(n):
        xs = (((s >> 18) ^ s) >> 27) & 0xFFFFFFFF
        rot = s >> 59
        out.append(((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF)
        s = (s * M + inc) & MASK
    return out


def main():
    assert [hex(x) for x in pcg32(6, 42, 54)] == PUBLISHED
    (HERE / "pcg32_kat.json").write_text(json.dumps(
        {"source": "PCG minimal C reference, pcg32-demo output, first block", "seed": 42, "seq": 54,
         "outputs": PUBLISHED}, indent=1) + "\n")
    cases = []
    for n, seed, seq in [(1000, 42, 54), (4097, 7, 54), (65536, 0x5EED, 7), (3, 0, 0), (10000, 2**63 + 5, 2**40)]:
        w = pcg32(n, seed, seq)
        cases.append({"n": n, "seed": seed, "seq": seq, "first8": w[:8],
                      "sha256": hashlib.sha256(struct.pack(f"<{n}I", *w)).hexdigest()})
    (HERE / "pcg32_streams.json").write_text(json.dumps({"generator": "make_pcg_fixtures.py (pure Python)",
                                                          "cases": cases}, indent=1) + "\n")


if __name__ == "__main__":
    main()
