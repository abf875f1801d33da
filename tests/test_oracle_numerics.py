"""Pins the oracle's numerics seam: csrmm, dense products, LU/det/inverse, LQ.

Known answers come from the reference's own self tests; everything else is
checked against LAPACK (through scipy) or numpy on the same inputs.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.linalg.lapack as lapack

import oracle_lib as orc

rng = np.random.default_rng(7)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def to_csr(D):
    rows, cols = np.nonzero(D)
    rp = np.zeros(D.shape[0] + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    return D[rows, cols].astype(np.complex128), cols.astype(np.int32), np.cumsum(rp).astype(np.int32)


# ---- csrmm: reference known answers ----------------------------------------
def test_csrmm_spblas_selftest():
    # src/Numerics/sp_blas.hpp:175-197: A(3,3)=1, A(2,1)=3, A(0,1)=9
    A = np.zeros((4, 4))
    A[3, 3], A[2, 1], A[0, 1] = 1.0, 3.0, 9.0
    B = np.array([[1., 2, 1, 5], [2, 5, 8, 7], [1, 8, 9, 9], [4, 1, 2, 3]], complex)
    val, col, rp = to_csr(A)
    Cn = np.zeros((4, 4), complex)
    orc.csrmm("N", 4, 4, 4, 1.0, val, col, rp, B, 0.0, Cn)
    assert np.array_equal(Cn.real, [[18, 45, 72, 63], [0, 0, 0, 0], [6, 15, 24, 21], [4, 1, 2, 3]])
    Ct = np.zeros((4, 4), complex)
    orc.csrmm("T", 4, 4, 4, 1.0, val, col, rp, B, 0.0, Ct)
    assert np.array_equal(Ct.real, [[0, 0, 0, 0], [12, 42, 36, 72], [0, 0, 0, 0], [4, 1, 2, 3]])


def test_csrmm_complex_2x2_selftest():
    # src/Numerics/tests/test_sparse_ops.cpp:226-272 (alpha = beta = 2)
    A = np.array([[2.0, 4.0], [0.0, 0.0]])
    val, col, rp = to_csr(A)
    B = np.array([[4.0, 3.0], [1.0, 5.0]], complex)
    Cm = np.array([[2.0, 1.3], [1.3, 1.3]], complex)
    orc.csrmm("N", 2, 2, 2, 2.0, val, col, rp, B, 2.0, Cm)
    np.testing.assert_allclose(Cm.ravel(), [28.0, 54.6, 2.6, 2.6], rtol=1e-14)


def test_csrmm_3x4_selftest():
    # src/Numerics/tests/test_sparse_ops.cpp:329-385
    A = np.zeros((3, 4))
    A[0, 0], A[0, 1], A[0, 3], A[1, 2], A[2, 1] = 2.0, 4.0, 1.0, 5.0, 1.0
    val, col, rp = to_csr(A)
    B = np.array([4, 3, 1, 5, 0, 0, 1, 0], float).reshape(4, 2).astype(complex)
    Cm = np.array([2.0, 1.3, 1.3, 1.3, 0, 0]).reshape(3, 2).astype(complex)
    orc.csrmm("N", 3, 2, 4, 2.0, val, col, rp, B, 2.0, Cm)
    np.testing.assert_allclose(Cm.ravel().real, [30.0, 54.6, 2.6, 2.6, 2.0, 10.0], rtol=1e-14)


@pytest.mark.parametrize("trans", ["N", "T", "H"])
def test_csrmm_random_vs_dense(trans):
    M, K, N = 13, 9, 5
    D = crand(M, K) * (rng.random((M, K)) < 0.4)
    val, col, rp = to_csr(D)
    alpha, beta = 0.3 - 1.1j, -0.7 + 0.2j
    if trans == "N":
        B = crand(K, N)
        C0 = crand(M, N)
        ref = alpha * D @ B + beta * C0
    else:
        B = crand(M, N)
        C0 = crand(K, N)
        op = D.T if trans == "T" else D.conj().T
        ref = alpha * op @ B + beta * C0
    Cm = C0.copy()
    orc.csrmm(trans, M, N, K, alpha, val, col, rp, B, beta, Cm)
    np.testing.assert_allclose(Cm, ref, rtol=1e-13, atol=1e-13)


def test_csrmm_skips_columns_beyond_K():
    # sparse.hpp:171: entries with col >= K are ignored
    D = crand(4, 6)
    val, col, rp = to_csr(D)
    B = crand(6, 3)
    Cm = np.zeros((4, 3), complex)
    orc.csrmm("N", 4, 3, 4, 1.0, val, col, rp, B, 0.0, Cm)
    np.testing.assert_allclose(Cm, D[:, :4] @ B[:4], rtol=1e-13)


# ---- dense products: ma_operations.hpp known answers ------------------------
def test_product_known_answers():
    # src/Numerics/ma_operations.hpp:445-489
    A = np.array([[1., 0, 1], [3, 5, 8], [4, 8, 9]])
    B = np.array([[6., 2, 8], [9, 5, 5], [1, 7, 9]])
    assert np.array_equal(orc.product("N", "N", A, B).real, [[7, 9, 17], [71, 87, 121], [105, 111, 153]])
    assert np.array_equal(orc.product("T", "N", A, B).real.ravel(), [37, 45, 59, 53, 81, 97, 87, 105, 129])
    assert np.array_equal(orc.product("N", "T", A, B).real.ravel(), [14, 14, 10, 92, 92, 110, 112, 121, 141])
    assert np.array_equal(orc.product("T", "T", A, B).real.ravel(), [44, 44, 58, 74, 65, 107, 94, 94, 138])
    # gemv known answers (ma_operations.hpp:326-410)
    M = np.array([[9., 24, 30, 2], [4, 10, 12, 1], [14, 16, 36, 20]])
    assert np.array_equal(orc.product("N", "N", M, np.array([[1.], [2], [3], [4]])).real.ravel(), [155, 64, 234])
    assert np.array_equal(orc.product("T", "N", M, np.array([[1.], [2], [3]])).real.ravel(), [59, 92, 162, 64])


def test_product_complex_ops():
    A, B = crand(5, 7), crand(7, 4)
    np.testing.assert_allclose(orc.product("N", "N", A, B), A @ B, rtol=1e-13)
    np.testing.assert_allclose(orc.product("H", "N", A.conj().T.copy(), B), A @ B, rtol=1e-13)
    a, b = 0.5 + 2j, -1 - 1j
    C0 = crand(5, 4)
    np.testing.assert_allclose(orc.product("N", "N", A, B, a, b, C0.copy()), a * A @ B + b * C0, rtol=1e-13)


# ---- LU / determinant / inverse ----------------------------------------------
def test_invert_selftest():
    # ma_operations.hpp:491-505: A*inv(A) = I at 1e-14
    A = np.array([[37., 45, 59], [53, 81, 97], [87, 105, 129]])
    Ai, det, info = orc.invert(A)
    assert info == 0
    np.testing.assert_allclose(Ai.real @ A, np.eye(3), atol=1e-14)
    np.testing.assert_allclose(det, np.linalg.det(A), rtol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 5, 16, 33])
def test_getrf_matches_lapack_pivots(n):
    A = crand(n, n)
    lu, piv, info = orc.getrf(A)
    # ma::getrf factorises the column-major view of the row-major array = A^T
    lu_ref, piv_ref, info_ref = lapack.zgetrf(A.T)
    assert info == 0 and info_ref == 0
    assert np.array_equal(piv, piv_ref + 1)
    np.testing.assert_allclose(lu.T, lu_ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("n", [2, 7, 16, 64])
def test_invert_and_det_vs_numpy(n):
    A = crand(n, n)
    Ai, det, info = orc.invert(A)
    assert info == 0
    np.testing.assert_allclose(Ai, np.linalg.inv(A), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(det, np.linalg.det(A), rtol=1e-11)
    d2, info = orc.determinant(A)
    np.testing.assert_allclose(d2, det, rtol=1e-14)


def test_det_closed_forms():
    # permutation matrices: det = sign of the permutation
    for perm in ([1, 0, 2], [2, 0, 1], [3, 2, 1, 0]):
        P = np.eye(len(perm))[perm].astype(complex)
        d, info = orc.determinant(P)
        assert d == np.linalg.det(P.real)
    # triangular: product of the diagonal
    U = np.triu(crand(6, 6)) + 3 * np.eye(6)
    d, _ = orc.determinant(U)
    np.testing.assert_allclose(d, np.prod(np.diag(U)), rtol=1e-13)
    # multiplicativity det(AB) = det(A) det(B)
    A, B = crand(8, 8), crand(8, 8)
    np.testing.assert_allclose(orc.determinant(A @ B)[0], orc.determinant(A)[0] * orc.determinant(B)[0], rtol=1e-11)


def test_singular_reports_info():
    A = crand(4, 4)
    A[:, 2] = 0
    d, info = orc.determinant(A)
    assert info > 0 and d == 0


# ---- LQ (zgelqf + zunglq on the column-major view) -------------------------
@pytest.mark.parametrize("M,N", [(8, 3), (64, 16), (40, 40)])
def test_lq_matches_lapack_qr(M, N):
    # LQ of A = W^T is the conjugate transpose of the QR of conj(W): both run
    # the same zlarfg reflectors, so R = L^H and Q_qr = Q_lq^H exactly.
    W = crand(M, N)
    Wq, Wl, tau = orc.lq(W)
    qr, tau_qr, _, info = lapack.zgeqrf(W.conj())
    assert info == 0
    np.testing.assert_allclose(tau[:N], tau_qr, rtol=1e-12, atol=1e-14)
    R = np.triu(qr[:N, :N])
    L = np.tril(Wl[:N, :N].T)  # column-major view of the gelqf output, lower part
    np.testing.assert_allclose(L, R.conj().T, rtol=1e-12, atol=1e-12)
    Q, _, info = lapack.zungqr(qr, tau_qr)
    np.testing.assert_allclose(Wq, Q[:, :N].conj(), rtol=1e-11, atol=1e-12)
    # orthonormality as the driver relies on it: W^T conj(W) = I
    np.testing.assert_allclose(Wq.T @ Wq.conj(), np.eye(N), atol=1e-13)
    # LAPACK's sign rule: real diagonal of L with sign opposite to Re(alpha)
    assert np.allclose(np.diag(L).imag, 0)


# ---- sparse ordering ---------------------------------------------------------
def test_compress_orders_rows_and_columns():
    nr = 7
    rows = rng.integers(0, nr - 2, 40).astype(np.int32)  # last rows empty
    cols = rng.permutation(200)[:40].astype(np.int32)
    vals = crand(40)
    r2, c2, v2, rp = orc.compress(nr, rows, cols, vals)
    key = r2.astype(np.int64) * 1000 + c2
    assert np.all(np.diff(key) > 0)
    for i in range(nr):
        assert rp[i] == np.searchsorted(r2, i) and rp[i + 1] == np.searchsorted(r2, i + 1)
    D = np.zeros((nr, 200), complex)
    D[rows, cols] = vals
    D2 = np.zeros((nr, 200), complex)
    D2[r2, c2] = v2
    assert np.array_equal(D, D2)
