"""File readers of the reference's inputs (ptc_read_tns / ptc_read_mtx):
1-based labels become 0-based, the order comes from the first data line,
dims follow readTensor / readMatrix (mtxTnsIO.cpp:5-63), comments and blank
lines are skipped, and malformed files are errors, not exits."""
import numpy as np
import pytest

import partc_amd as P


def test_tns_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    h = np.unique(rng.integers(0, [7, 5, 9, 3], size=(200, 4)).astype(np.uint32), axis=0)
    v = rng.uniform(0, 1, h.shape[0])
    T = P.Tensor(h, v)
    f = tmp_path / "t.tns"
    T.write_tns(f)
    R = P.read_tns(f)
    assert R.d == 4 and R.N == h.shape[0]
    assert np.array_equal(R.hedges_array, h)
    assert np.array_equal(R.val_array, v)  # 17 significant digits round-trip exactly
    assert np.array_equal(R.dimensions_array, h.max(axis=0) + 1)  # largest 1-based label


def test_tns_comments_and_order(tmp_path):
    f = tmp_path / "c.tns"
    f.write_text("# a comment\n\n1 2 3 0.5\n2 1 1 1.25e-3\n   3 3 4   -2\n")
    R = P.read_tns(f)
    assert R.d == 3 and R.N == 3
    assert R.hedges_array.tolist() == [[0, 1, 2], [1, 0, 0], [2, 2, 3]]
    assert R.val_array.tolist() == [0.5, 1.25e-3, -2.0]
    assert R.dimensions_array.tolist() == [3, 3, 4]


@pytest.mark.parametrize("text", ["1 2 0.5\n1 2 3 0.5\n", "0 1 0.5\n", "1 x 0.5\n", "1 2\n2 3 4\n"])
def test_tns_malformed(tmp_path, text):
    f = tmp_path / "bad.tns"
    f.write_text(text)
    with pytest.raises(P.PtcError):
        P.read_tns(f)


def test_mtx(tmp_path):
    f = tmp_path / "m.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n4 5 3\n1 1 2.5\n4 2 -1\n2 5 7e-1\n")
    R = P.read_mtx(f)
    assert R.d == 2 and R.N == 3
    assert R.hedges_array.tolist() == [[0, 0], [3, 1], [1, 4]]
    assert R.val_array.tolist() == [2.5, -1.0, 0.7]
    assert R.dimensions_array.tolist() == [4, 5]


def test_mtx_pattern_and_errors(tmp_path):
    f = tmp_path / "p.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 2\n3 3\n")
    R = P.read_mtx(f)
    assert R.val_array.tolist() == [1.0, 1.0]
    for text, code in (("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -6),
                       ("%%MatrixMarket matrix coordinate real general\n3 3 3\n1 1 1\n2 2 2\n", -7),
                       ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", -5)):
        g = tmp_path / "e.mtx"
        g.write_text(text)
        with pytest.raises(P.PtcError) as ei:
            P.read_mtx(g)
        assert ei.value.code == code
    with pytest.raises(P.PtcError):
        P.read_tns(tmp_path / "missing.tns")
