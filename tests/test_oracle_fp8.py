"""Pins the oracle's restatement of the FP8 index path (or_fp8_encode /
or_fp8_decode / or_fp8_params / or_fp8_codes / or_fp8_brute_batch) before
the GPU is checked against it.

* Encoder known answers worked by hand: zero, one, the largest finite value
  448, saturation, sign, the smallest normal and subnormal, ties to even in
  the subnormal and normal ranges, a carry into the next binade.
* Decode of all 256 codes against the E4M3 definition written out
  independently here (bias 7, 3 mantissa bits, subnormals m x 2^-9).
* The encoder returns the nearest code (ties: the even mantissa) by
  exhaustive search over the 253 finite non-negative values, on random and
  on every midpoint.
* Parameters: midrange centres and 448 / largest half-range.
* Reduction to the exact search: with over x k at least the passing rows,
  the FP8 batch equals the exact brute force bit for bit; recall never
  drops as the overfetch grows.
"""
import numpy as np

import oracle as O
from paves_amd import engine as E
from paves_amd import synth


def _spec_value(c):
    """E4M3 (e4m3fn) code -> value, from the format's definition."""
    s = -1.0 if c & 0x80 else 1.0
    e = (c >> 3) & 0xF
    m = c & 0x7
    if e == 0:
        return s * m / 8.0 * 2.0 ** (1 - 7)
    return s * (1.0 + m / 8.0) * 2.0 ** (e - 7)


def test_decode_table_matches_definition():
    t = O.fp8_decode_table()
    for c in range(256):
        assert t[c] == np.float32(_spec_value(c)), hex(c)
    assert t[0x7E] == 448.0 and t[0x08] == 2.0 ** -6 and t[0x01] == 2.0 ** -9


def test_encoder_known_answers():
    kat = {
        0.0: 0x00, 1.0: 0x38, -2.0: 0xC0, 448.0: 0x7E, 1000.0: 0x7E, -1e9: 0xFE,
        2.0 ** -6: 0x08, 2.0 ** -9: 0x01,
        2.0 ** -10: 0x00,          # subnormal tie 0.5 -> 0 (even), no -0
        -2.0 ** -10: 0x00,
        3 * 2.0 ** -10: 0x02,      # 1.5 units -> 2 (even)
        5 * 2.0 ** -10: 0x02,      # 2.5 units -> 2 (even)
        1.0625: 0x38,              # 1 + 1/16: tie between 1.0 and 1.125 -> 1.0
        1.1875: 0x3A,              # 1 + 1.5/8 -> mantissa 2 (even)
        15.5: 0x58,                # 1.9375 x 8: mantissa 7.5 -> 8 -> carry to 16
        440.0: 0x7E,               # between 416 and 448 above the midpoint 432
        432.0: 0x7E,               # tie 416 / 448 -> mantissa 6 (even) = 448
        424.0: 0x7D,               # below the midpoint -> 416
        -0.3: 0xAA,                # 0.3 = 1.2 x 2^-2: mantissa 1.6 -> 2 -> 0x2A
    }
    for v, code in kat.items():
        assert O.fp8_encode(v) == code, (v, hex(O.fp8_encode(v)), hex(code))


def test_encoder_is_nearest_even():
    t = O.fp8_decode_table().astype(np.float64)
    pos = np.array([c for c in range(0x7F)])           # 0x00..0x7E, finite, >= 0
    vals = t[pos]
    rng = np.random.default_rng(11)
    xs = list(rng.uniform(0, 448, 3000)) + list(np.exp(rng.uniform(-9, 6, 3000)))
    # every midpoint between neighbouring codes (a tie)
    xs += list((vals[1:] + vals[:-1]) / 2.0)
    for x in xs:
        x = float(np.float32(x))
        err = np.abs(vals - x)
        best = np.flatnonzero(err == err.min())
        if len(best) > 1:
            best = [b for b in best if (pos[b] & 1) == 0]
        want = int(pos[best[0]])
        got = O.fp8_encode(x)
        assert got == want, (x, hex(got), hex(want))
        assert O.fp8_encode(-x) == (want | 0x80 if want else 0)


def test_params_and_codes():
    v = np.array([[0.0, 10.0, 3.0], [2.0, -10.0, 3.0], [1.5, 5.0, 3.0]], np.float32)
    mid, sc = O.fp8_params(v)
    assert mid.tolist() == [1.0, 0.0, 3.0]
    assert sc == np.float32(448.0) / np.float32(10.0)
    c = O.fp8_codes(v, mid, sc)
    # (0-1)*44.8 = -44.8 = -1.4 x 2^5: exponent field 5 + 7 = 12, mantissa
    # 0.4 x 8 = 3.2 -> 3: 0x63 with the sign 0xE3; 10*44.8 = 448 -> 0x7E;
    # the constant dim -> 0
    assert c[0].tolist() == [0xE3, 0x7E, 0x00]
    assert c[1].tolist() == [0x63, 0xFE, 0x00]


def _data(n=3000, nq=24, dim=48, seed=5):
    d = synth.gmm768(n, nq, seed=seed, dim=dim, ncent=32)
    cols = synth.columns(d)
    rows, vals = synth.postings(cols)
    post = E.Postings(rows, vals)
    preds = [synth.filter_pred(synth.FILTER_CLASSES[i % 6], d["qcent"][i]) for i in range(nq)]
    b = E.concat([E.serialize(p, n, post) for p in preds], n)
    return d, cols, b


def test_full_overfetch_equals_exact_and_recall_monotone():
    d, cols, b = _data()
    base, qs = d["base"], d["queries"]
    k = 10
    oh, oc, ot, passing = O.brute_batch(base, cols, qs, b.qd, b.nodes, b.children, b.invals, k)
    # over x k >= every query's passing rows (<= 3000 <= 4096): exact
    fh, fc, ft = O.fp8_brute_batch(base, cols, qs, b.qd, b.nodes, b.children, b.invals, k, 400)
    assert np.array_equal(fc, oc) and np.array_equal(ft, ot)
    for q in range(len(qs)):
        c = int(oc[q])
        assert np.array_equal(fh[q, :c]["node"], oh[q, :c]["node"])
        assert np.array_equal(fh[q, :c]["dist"].view(np.uint32), oh[q, :c]["dist"].view(np.uint32))
    prev = -1.0
    for over in (1, 2, 4, 8):
        fh, fc, _ = O.fp8_brute_batch(base, cols, qs, b.qd, b.nodes, b.children, b.invals, k, over)
        hit = sum(len(set(fh[q, :fc[q]]["node"].tolist()) & set(oh[q, :oc[q]]["node"].tolist()))
                  for q in range(len(qs)))
        rec = hit / max(1, int(oc.sum()))
        assert rec >= prev
        prev = rec
    assert prev >= 0.95
