"""Explicit-matrix operators (operators.py:11-89) on the CPU: with the
reference's ./matrixes/A_<basis>_sparse_<sizes>.npz present, the operator
multiplies by that CSR matrix (matvec, rmatvec, adjoint view), loads it
without unpickling anything, and rejects a file of the wrong shape."""
import numpy as np
import pytest


def _random_csr(shape, seed):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    return sp.random(shape[0], shape[1], density=0.05, random_state=rng, format="csr")


@pytest.mark.parametrize("cls, name, dims", [("A_B0_operator_matrix", "B0", (6, 5, 4, 9)),
                                             ("A_B1_operator_matrix", "B1", (6, 5, 4, 9)),
                                             ("A_H0_operator_matrix", "H0", (7, 4, 3, 8)),
                                             ("A_Hl_operator_matrix", "Hl", (7, 4, 3, 8))])
def test_matrix_operator_loads_saved_matrix(tmp_path, cls, name, dims):
    import scipy.sparse as sp
    from sgd_amd import operators as ops
    d0, d1, nnp, nu = dims
    M = _random_csr((nnp * nu, d0 * d1), 3)
    sp.save_npz(str(tmp_path / ("A_%s_sparse_%d_%d_%d_%d.npz" % ((name,) + dims))), M)
    args = (d0, d1, nnp, nu, 1.0, 1.0, 1.2, 200.0, 100.0) if name in ("B0", "B1") else \
        (d0, d1, nnp, nu, 1.0, 1.2, 200.0, 100.0)
    op = getattr(ops, cls)(*args, matrix_dir=str(tmp_path))
    assert op.shape == (nnp * nu, d0 * d1) and op.source.endswith(".npz")
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(op.shape[1]), rng.standard_normal(op.shape[0])
    assert np.allclose(op.matvec(x), M @ x) and np.allclose(op.rmatvec(y), M.T @ y)
    assert np.allclose(op.H.matvec(y), M.T @ y)
    bad = _random_csr((nnp * nu + 1, d0 * d1), 4)
    sp.save_npz(str(tmp_path / ("A_%s_sparse_%d_%d_%d_%d.npz" % ((name,) + dims))), bad)
    with pytest.raises(ValueError):
        getattr(ops, cls)(*args, matrix_dir=str(tmp_path))
