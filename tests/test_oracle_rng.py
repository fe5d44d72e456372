"""Pins the oracle's random streams (math/random.hpp:24-120).

* PCG32 core: the published pcg32 demo stream, pcg32_srandom_r(42, 54).
* SplitMix64 / MurmurHash3 fmix32: published reference outputs.
* seed_random_xyz: the reference seeds PCG the pcg32_srandom_r way with
  initstate = splitmix64(seed2), initseq = splitmix64(seed1), then discards 8.
* randf: closed form (u32 >> 9) * 2^-23.
* Statistical properties ported from test_random.cpp (range, mean, variance,
  chi-square, bit balance, serial correlation, distinct seeds).
"""

import ctypes as C

import numpy as np

import oracle

L = oracle.lib()
M64 = (1 << 64) - 1
MUL = 6364136223846793005


def _stream(state, inc, n):
    out = (C.c_uint32 * n)()
    L.orc_pcg32_stream(state, inc, n, out)
    return list(out)


def _srandom(initstate, initseq):
    inc = ((initseq << 1) | 1) & M64
    s = (0 * MUL + inc) & M64
    s = (s + initstate) & M64
    s = (s * MUL + inc) & M64
    return s, inc


def test_pcg32_published_demo_stream():
    s, inc = _srandom(42, 54)
    assert _stream(s, inc, 6) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_splitmix64_published_outputs():
    g = 0x9E3779B97F4A7C15
    # SplitMix64 seeded with 0: first three outputs of the reference generator
    assert L.orc_splitmix64(0) == 0xE220A8397B1DCDAF
    assert L.orc_splitmix64(g) == 0x6E789E6AA1B965F4
    assert L.orc_splitmix64((2 * g) & M64) == 0x06C45D188009454F


def test_fmix32_published_outputs():
    assert L.orc_fmix32(0) == 0
    assert L.orc_fmix32(1) == 0x514E28B7


def test_seed_xyz_is_pcg_srandom_plus_8_discards():
    for (x, y, s) in [(0, 0, 0), (5, 7, 3), (1023, 511, 64), (4000000000, 1, 77)]:
        h1 = L.orc_jenkins((x + 1) & 0xFFFFFFFF)
        h2 = L.orc_jenkins((y + 1 + h1) & 0xFFFFFFFF)
        h3 = L.orc_jenkins((s + 1 + h2) & 0xFFFFFFFF)
        seed1 = (L.orc_fmix32(h1 ^ h3) << 32) | L.orc_fmix32(h2)
        seed2 = (L.orc_fmix32(h2 ^ h1) << 32) | L.orc_fmix32(h3)
        st, inc = _srandom(L.orc_splitmix64(seed2), L.orc_splitmix64(seed1))
        for _ in range(8):
            st = (st * MUL + inc) & M64
        a, b = C.c_uint64(), C.c_uint64()
        L.orc_seed(x, y, s, C.byref(a), C.byref(b))
        assert (a.value, b.value) == (st, inc)


def test_randf_closed_form():
    a, b = C.c_uint64(), C.c_uint64()
    L.orc_seed(3, 4, 5, C.byref(a), C.byref(b))
    u = _stream(a.value, b.value, 64)
    f = (C.c_float * 64)()
    L.orc_randf_stream(3, 4, 5, 64, f)
    want = np.array([(x >> 9) for x in u], np.float64) * 2.0 ** -23
    np.testing.assert_array_equal(np.array(list(f), np.float32), want.astype(np.float32))


def _many(n_pixels=4096, per=16):
    f = (C.c_float * per)()
    vals = []
    for i in range(n_pixels):
        L.orc_randf_stream(i % 64, i // 64, 1, per, f)
        vals.extend(f)
    return np.array(vals, np.float64)


def test_statistics_like_reference():
    v = _many()
    assert v.min() >= 0.0 and v.max() < 1.0                      # RandfInRange
    assert abs(v.mean() - 0.5) < 0.01                              # MeanApproximately0_5
    assert abs(v.var() - 1.0 / 12.0) < 0.005                       # VarianceCorrect
    h, _ = np.histogram(v, bins=10, range=(0, 1))
    e = len(v) / 10
    chi2 = ((h - e) ** 2 / e).sum()
    assert 1.0 < chi2 < 30.0                                       # ChiSquareTest
    ac = np.corrcoef(v[:-1], v[1:])[0, 1]
    assert abs(ac) < 0.03                                          # NoObviousSerialCorrelation


def test_bits_balanced_and_seeds_distinct():
    a, b = C.c_uint64(), C.c_uint64()
    L.orc_seed(0, 0, 0, C.byref(a), C.byref(b))
    u = np.array(_stream(a.value, b.value, 20000), np.uint64)
    ratios = [((u >> k) & 1).mean() for k in range(32)]
    assert all(abs(r - 0.5) < 0.02 for r in ratios)                # EachBitApproximately50Percent
    firsts = set()
    for x in range(8):
        for y in range(8):
            L.orc_seed(x, y, 0, C.byref(a), C.byref(b))
            firsts.add(_stream(a.value, b.value, 1)[0])
    assert len(firsts) == 64                                       # XYZSeeding_DifferentPixels


def test_jenkins_matches_reference_vectors():
    """jenkins_hash (random.hpp:49-57) on 328 inputs: the fixture was
    evaluated from the reference's own statements, parsed from its text by
    tests/golden/make_jenkins_vectors.py."""
    import json
    import os

    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "jenkins32_vectors.json")))
    assert len(fx["statements"]) == 6 and len(fx["pairs"]) > 300
    bad = [(a, b) for a, b in fx["pairs"] if L.orc_jenkins(a) != b]
    assert not bad, f"{len(bad)} mismatches, first {bad[:3]}"
