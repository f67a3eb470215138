"""The product's MatrixMarket reader follows MMReader (file order per row,
b = -rhs); checked against the oracle's C reader on the golden files."""
import os

import numpy as np

import oracle_lib as O
from fvm_amd import read_matrix_market

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_reader_matches_oracle_on_goldens():
    for mat, rhs in [("MatrixMarket226.dat", "rhs226.dat"), ("cav32_matrix_mesh0.mat", "cav32_matrix.rhs")]:
        ls = read_matrix_market(os.path.join(G, mat), os.path.join(G, rhs))
        m = O.load_mm(os.path.join(G, mat), os.path.join(G, rhs))
        assert ls.n_self == m["n_self"]
        assert np.array_equal(ls.row, m["row"])
        assert np.array_equal(ls.col, m["col"])
        assert np.array_equal(ls.off, m["off"])
        assert np.array_equal(ls.diag, m["diag"])
        assert np.array_equal(ls.b, m["b"])


def test_symmetric_reader_mirrors_entries(tmp_path):
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 4\n1 1 -2\n2 1 1\n2 2 -2\n3 2 0.5\n")
    r = tmp_path / "r.txt"
    r.write_text("1\n2\n3\n")
    ls = read_matrix_market(str(p), str(r))
    assert ls.isSymmetric
    assert ls.row.tolist() == [0, 1, 3, 4]
    assert ls.col.tolist() == [1, 0, 2, 1]
    assert ls.off.tolist() == [1.0, 1.0, 0.5, 0.5]
    assert ls.b.tolist() == [-1.0, -2.0, -3.0]
