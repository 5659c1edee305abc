"""Host side of the format converters: the Matrix Market reader against
scipy.io.mmread on the fixture files and on generated files of every field /
symmetry kind it accepts."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _dense(rows, cols, r, c, v):
    D = np.zeros((rows, cols), np.float64)
    np.add.at(D, (r, c), v)
    return D


@pytest.mark.parametrize("name", ["1138_bus.mtx", "iris_dataset_30NN.mtx"])
def test_reader_matches_scipy_on_fixtures(bws, name):
    path = os.path.join(GOLDEN, name)
    rows, cols, r, c, v = bws.read_matrix_market(path)
    ref = scipy.io.mmread(path).toarray()
    assert (rows, cols) == ref.shape
    assert np.allclose(_dense(rows, cols, r, c, v), ref.astype(np.float32), rtol=0, atol=0)


@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
def test_reader_kinds(bws, tmp_path, rng, field, symmetry):
    n = 12
    A = sp.random(n, n, density=0.3, random_state=int(rng.integers(1 << 30)), format="coo")
    if field == "integer":
        A.data = np.round(A.data * 10) + 1
    if symmetry != "general":
        L = sp.tril(A, k=-1 if symmetry == "skew-symmetric" else 0).tocoo()
        if symmetry == "symmetric":
            A = (L + sp.tril(L, k=-1).T).tocoo()
        else:
            A = (L - L.T).tocoo()
    path = str(tmp_path / "m.mtx")
    if field == "pattern":
        A.data = np.ones_like(A.data)
    scipy.io.mmwrite(path, A, field=field if field != "real" else None, symmetry=symmetry)
    rows, cols, r, c, v = bws.read_matrix_market(path)
    ref = scipy.io.mmread(path).toarray()
    assert np.array_equal(_dense(rows, cols, r, c, v), ref.astype(np.float32).astype(np.float64))


def test_reader_rejects_array_format(bws, tmp_path):
    path = str(tmp_path / "a.mtx")
    scipy.io.mmwrite(path, np.ones((2, 2)))
    with pytest.raises(ValueError):
        bws.read_matrix_market(path)
