#!/usr/bin/env python3
"""Writes tests/golden/known_answers.json: small fields with their packed
binary maps and 2x2x2 pattern histograms, computed by plain Python loops from
the definitions alone (no numpy arithmetic, no oracle, no library):

* threshold bit (src/binMapCUDA.cuh:131-147): d = 1 - (f < t) + (f > t)
  (d = 1 when unordered, i.e. NaN), bit = (m >> d) & 1; a float32 field is
  compared with t rounded to float32;
* PBM (src/binMapUtil.c:60-87): bit x & 31 of word (x >> 5) + y rowWS +
  z rowWS Y, rowWS = ceil(X / 32), tail bits zero;
* cell code (src/bpfdCUDA.cuh:45-61): bit i = P(x + (i & 1), y + ((i >> 1) & 1),
  z + (i >> 2)) over the (X-1)(Y-1)(Z-1) cells.

The reference holds no test vectors of its own; these fixtures pin the
oracle and the GPU kernels to the definitions on cases small enough to
enumerate, including ragged rows, dims of 2, special values and all eight
relation masks. Run from the repository root: python tests/golden/make_fixtures.py
"""
import json
import math
import os
import random
import struct

HERE = os.path.dirname(os.path.abspath(__file__))

RELATIONS = {"<": 1, "=": 2, "<=": 3, ">": 4, "!=": 5, ">=": 6, "": 0, "!": 7}


def to_f32(v):
    return struct.unpack("<f", struct.pack("<f", v))[0]


def bit(f, t, m):
    if math.isnan(f) or math.isnan(t):
        d = 1
    else:
        d = 1 - (1 if f < t else 0) + (1 if f > t else 0)
    return (m >> d) & 1


def pbm_words(field, X, Y, Z, t, m):
    rowWS = (X + 31) // 32
    words = [0] * (rowWS * Y * Z)
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                if bit(field[(z * Y + y) * X + x], t, m):
                    words[(x >> 5) + y * rowWS + z * rowWS * Y] |= 1 << (x & 31)
    return words, rowWS


def histogram(words, X, Y, Z, rowWS):
    def P(x, y, z):
        return (words[(x >> 5) + y * rowWS + z * rowWS * Y] >> (x & 31)) & 1

    h = [0] * 256
    for z in range(Z - 1):
        for y in range(Y - 1):
            for x in range(X - 1):
                c = 0
                for i in range(8):
                    c |= P(x + (i & 1), y + ((i >> 1) & 1), z + (i >> 2)) << i
                h[c] += 1
    return h


def enc(v):
    """JSON has no NaN/inf: special values are strings."""
    if math.isnan(v):
        return "nan"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return v


def make_cases():
    rng = random.Random(20240917)
    specials = [0.0, -0.0, 1e-40, -1e-40, 5e-324, float("nan"), float("inf"), float("-inf"), 0.25, -0.25, 1.0]
    cases = []
    shapes = [(2, 2, 2), (3, 2, 2), (33, 3, 2), (31, 2, 3), (5, 4, 3), (65, 2, 2), (40, 3, 3), (7, 6, 5)]
    for i, (X, Y, Z) in enumerate(shapes):
        for dtype in ("f32", "f64"):
            n = X * Y * Z
            if i % 3 == 0:
                vals = [rng.choice(specials) for _ in range(n)]
            else:
                vals = [rng.uniform(-1.0, 1.0) for _ in range(n)]
            if dtype == "f32":
                vals = [to_f32(v) for v in vals]
            rel = list(RELATIONS)[(i * 2 + (dtype == "f64")) % len(RELATIONS)]
            t_in = [0.0, 0.25, -0.25, 0.1][i % 4]
            t = to_f32(t_in)  # setBinMapF64 stores the float threshold
            if dtype == "f32":
                t_cmp = t
            else:
                t_cmp = t  # the double field is compared with the (float-valued) double
            m = RELATIONS[rel]
            words, rowWS = pbm_words(vals, X, Y, Z, t_cmp, m)
            h = histogram(words, X, Y, Z, rowWS)
            cases.append({"dims": [X, Y, Z], "dtype": dtype, "relop": rel, "t": t_in, "mask": m,
                          "field": [enc(v) for v in vals], "pbm": words, "hist": h})
    return cases


def main():
    cases = make_cases()
    with open(os.path.join(HERE, "known_answers.json"), "w") as fh:
        json.dump({"generator": "tests/golden/make_fixtures.py", "cases": cases}, fh, separators=(",", ":"))
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
