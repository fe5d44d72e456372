"""Writes tests/golden/jenkins32_vectors.json: input/output pairs of the
reference's 32-bit Jenkins integer hash (LucidRenderer/cpp_renderer/include/
math/random.hpp:49-57, `jenkins_hash`, the first stage of seed_random_xyz).

The hash has no published test vectors, so the vectors come from the
reference's own text: this script reads random.hpp as text, parses each of
the six mixing statements of `jenkins_hash`, which all have the form

    a = (a <op1> <constant>) <op2> (a <shift> <amount>);

and evaluates that statement sequence in Python integer arithmetic modulo
2^32.  Nothing of the reference is imported or executed.  It runs in the
build container (where the reference is mounted); the tests read only the
JSON it writes.

    python tests/golden/make_jenkins_vectors.py [/root/reference]
"""

from __future__ import annotations

import json
import os
import re
import sys

M32 = 0xFFFFFFFF
STMT = re.compile(r"a\s*=\s*\(\s*a\s*([+^])\s*(0x[0-9a-fA-F]+)\s*\)\s*([+^])\s*\(\s*a\s*(<<|>>)\s*(\d+)\s*\)\s*;")


def parse(ref_root: str):
    path = os.path.join(ref_root, "LucidRenderer", "cpp_renderer", "include", "math", "random.hpp")
    text = open(path, encoding="utf-8").read()
    m = re.search(r"inline\s+uint32_t\s+jenkins_hash\s*\(\s*uint32_t\s+a\s*\)\s*\{(.*?)return\s+a\s*;", text, re.S)
    if not m:
        raise SystemExit("jenkins_hash not found in random.hpp")
    steps = [(op1, int(c, 16), op2, sh, int(n)) for op1, c, op2, sh, n in STMT.findall(m.group(1))]
    if len(steps) != 6:
        raise SystemExit(f"expected 6 mixing statements, parsed {len(steps)}")
    return steps, os.path.relpath(path, ref_root)


def apply(steps, a: int) -> int:
    for op1, c, op2, sh, n in steps:
        left = (a + c) & M32 if op1 == "+" else a ^ c
        right = (a << n) & M32 if sh == "<<" else a >> n
        a = (left + right) & M32 if op2 == "+" else left ^ right
    return a


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    steps, src = parse(ref_root)
    # inputs: the small keys seed_random_xyz hashes (x + 1, y + 1 + h1, ...),
    # powers of two, all-ones patterns, and a seeded spread over 2^32
    inputs = list(range(0, 64)) + [1 << k for k in range(32)] + [(1 << k) - 1 for k in range(1, 33)]
    x = 0x12345678
    for _ in range(200):
        x = (x * 1664525 + 1013904223) & M32
        inputs.append(x)
    out = {"source": f"{src}: jenkins_hash", "statements": [list(s) for s in steps],
           "pairs": [[a, apply(steps, a)] for a in inputs]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jenkins32_vectors.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")
    print(f"wrote {len(out['pairs'])} vectors to {dst}")


if __name__ == "__main__":
    main()
