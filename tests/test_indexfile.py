"""The PAVES index-file reader and writer (paves_amd/indexfile.py, format of
engine.cpp:190-215) against files the reference engine itself writes and
opens: a reference-built file (vectors, attributes with NULLs, postings,
HNSW sections) read here equals what the oracle's independent parser reads
and what was put in; a file written here opens in the reference engine and
its BruteCursor answers as on the reference-built file."""
import numpy as np
import pytest

import oracle as O
from conftest import have_reference_build
from paves_amd import engine as E
from paves_amd import indexfile, synth

pytestmark = pytest.mark.skipif(not have_reference_build(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def refidx(tmp_path_factory):
    d = synth.gmm128(3000, 8, seed=71)
    cols = synth.columns(d)
    cols[1, ::37] = np.nan
    path = str(tmp_path_factory.mktemp("idx") / "ref.idx")
    O.RefEngine().build(path, d["base"], cols, M=8, efc=40, threads=1)
    return d, cols, path


def test_read_reference_written_file(refidx):
    d, cols, path = refidx
    f = indexfile.read(path)
    o = O.read_index_file(path)
    assert f["count"] == 3000 and f["dim"] == 128 and f["ncols"] == 2 and f["M"] == 8
    assert f["l0_stride"] == 17 and f["entry_node"] == o["entry_node"]
    assert f["max_level"] == o["max_level"]
    for k in ("vectors", "tids", "post_rows", "levels", "l0", "upper_offs", "upper"):
        assert np.array_equal(f[k], o[k]), k
    for k in ("cols", "post_vals"):
        assert np.array_equal(f[k], o[k], equal_nan=True), k
    assert np.array_equal(f["vectors"], d["base"])
    assert np.array_equal(f["cols"], cols, equal_nan=True)
    assert np.array_equal(f["tids"], np.arange(3000, dtype=np.uint64))
    # postings: every column's rows sorted by value (ties by row), NULLs
    # last; the reference leaves the order among NULL rows open, so only
    # their set is compared
    rows, vals = synth.postings(cols)
    for c in range(2):
        nn = int((~np.isnan(cols[c])).sum())
        assert np.array_equal(f["post_rows"][c, :nn], rows[c, :nn]), c
        assert np.array_equal(f["post_vals"][c, :nn], vals[c, :nn]), c
        assert set(f["post_rows"][c, nn:].tolist()) == set(rows[c, nn:].tolist()), c
        assert np.isnan(f["post_vals"][c, nn:]).all(), c


def test_written_file_opens_in_reference(refidx, tmp_path):
    d, cols, path = refidx
    f = indexfile.read(path)
    out = str(tmp_path / "mine.idx")
    indexfile.write(out, f["vectors"], f["cols"], M=f["M"], efc=f["efc"], levels=f["levels"],
                    l0=f["l0"], upper_offs=f["upper_offs"], upper=f["upper"],
                    entry_node=f["entry_node"], max_level=f["max_level"],
                    post_rows=f["post_rows"], post_vals=f["post_vals"])
    with open(path, "rb") as a, open(out, "rb") as b:
        assert a.read() == b.read()
    ref = O.RefEngine()
    h1, h2 = ref.open(path), ref.open(out)
    try:
        for qi in range(4):
            p = synth.filter_pred("price<500", d["qcent"][qi])
            p1, k1 = ref.compile(h1, p)
            p2, k2 = ref.compile(h2, p)
            assert ref.brute(h1, d["queries"][qi], p1, 10) == ref.brute(h2, d["queries"][qi], p2, 10)
            ref.L.fv_pred_free(p1)
            ref.L.fv_pred_free(p2)
    finally:
        ref.close(h1)
        ref.close(h2)
