"""Matrix Market input (spx_mm_read): hand-traced small files for every
field/symmetry the reference's loader handles (main.cu:164-319), the error
codes, a round trip of a generated matrix (bit-exact CSR), and random files
with duplicates against a dense accumulation.  Host code: runs without a GPU."""
import numpy as np
import pytest

import spx
from spx import workloads as W


def write(tmp_path, text, name="a.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_general_real_keeps_file_order_within_rows(tmp_path):
    path = write(tmp_path, """%%MatrixMarket matrix coordinate real general
% a comment

3 4 5
3 1 -2.5
1 4 1e-3
1 2 7
3 3 0.5
2 2 4
""")
    A = spx.read_mtx(path)
    assert (A["m"], A["n"], A["nnz"]) == (3, 4, 5)
    assert A["rowptr"].tolist() == [0, 2, 3, 5]
    assert A["col"].tolist() == [3, 1, 1, 0, 2]
    assert A["val"].tolist() == [1e-3, 7.0, 4.0, -2.5, 0.5]


def test_pattern_reads_ones_and_integer_reads_doubles(tmp_path):
    A = spx.read_mtx(write(tmp_path, "%%MatrixMarket matrix coordinate pattern general\n2 2 3\n1 1\n2 1\n2 2\n"))
    assert A["val"].tolist() == [1.0, 1.0, 1.0] and A["col"].tolist() == [0, 0, 1]
    B = spx.read_mtx(write(tmp_path, "%%MatrixMarket matrix coordinate integer general\n1 3 2\n1 3 -7\n1 1 12\n",
                           "b.mtx"))
    assert B["val"].tolist() == [-7.0, 12.0] and B["col"].tolist() == [2, 0]


@pytest.mark.parametrize("sym,sign", [("symmetric", 1.0), ("hermitian", 1.0), ("skew-symmetric", -1.0)])
def test_symmetric_files_mirror_off_diagonal_entries(tmp_path, sym, sign):
    diag = "" if sym == "skew-symmetric" else "2 2 5\n"
    n_ent = 2 if sym == "skew-symmetric" else 3
    path = write(tmp_path, f"%%MatrixMarket matrix coordinate real {sym}\n3 3 {n_ent}\n2 1 1.5\n{diag}3 1 -4\n")
    A = spx.read_mtx(path)
    dense = np.zeros((3, 3))
    for r in range(3):
        for k in range(A["rowptr"][r], A["rowptr"][r + 1]):
            dense[r, A["col"][k]] += A["val"][k]
    exp = np.zeros((3, 3))
    exp[1, 0], exp[2, 0] = 1.5, -4.0
    exp[0, 1], exp[0, 2] = sign * 1.5, sign * -4.0
    if diag:
        exp[1, 1] = 5.0
    assert np.array_equal(dense, exp)
    # row 0 holds the mirrors in file order
    assert A["col"][A["rowptr"][0]:A["rowptr"][1]].tolist() == [1, 2]


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", -3),
    ("%%MatrixMarket matrix array real general\n1 1\n1\n", -2),
    ("%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", -2),
    ("%%MatrixMarket matrix coordinate real general\n1 1\n1 1 1\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", -5),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", -5),
])
def test_errors(tmp_path, text, code):
    with pytest.raises(spx.SpxError) as e:
        spx.read_mtx(write(tmp_path, text))
    assert e.value.code == code
    with pytest.raises(spx.SpxError) as e:
        spx.read_mtx(str(tmp_path / "missing.mtx"))
    assert e.value.code == -1


def test_round_trip_of_a_generated_matrix_is_bit_exact(tmp_path):
    A = W.zipf(3000, 1.4, 200, seed=4)
    rows = np.repeat(np.arange(A["m"]), np.diff(A["rowptr"]))
    lines = [f"%%MatrixMarket matrix coordinate real general\n{A['m']} {A['n']} {A['nnz']}\n"]
    lines += [f"{r + 1} {c + 1} {float(v)!r}\n" for r, c, v in zip(rows, A["col"], A["val"])]
    B = spx.read_mtx(write(tmp_path, "".join(lines)))
    assert np.array_equal(B["rowptr"], A["rowptr"])
    assert np.array_equal(B["col"], A["col"])
    assert np.array_equal(B["val"], A["val"])


def test_random_files_with_duplicates_match_dense_accumulation(tmp_path):
    rng = np.random.default_rng(5)
    for t in range(5):
        m, n, L = int(rng.integers(1, 40)), int(rng.integers(1, 40)), int(rng.integers(0, 200))
        r, c = rng.integers(0, m, L), rng.integers(0, n, L)
        v = rng.integers(-9, 10, L).astype(float)
        text = f"%%MatrixMarket matrix coordinate real general\n{m} {n} {L}\n"
        text += "".join(f"{a + 1} {b + 1} {x}\n" for a, b, x in zip(r, c, v))
        A = spx.read_mtx(write(tmp_path, text, f"r{t}.mtx"))
        dense = np.zeros((m, n))
        np.add.at(dense, (r, c), v)
        got = np.zeros((m, n))
        rows = np.repeat(np.arange(m), np.diff(A["rowptr"]))
        np.add.at(got, (rows, A["col"]), A["val"])
        assert np.array_equal(got, dense)
        assert A["nnz"] == L
