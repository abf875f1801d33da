"""numpy binding of oracle/liboracle.so (test infrastructure: the checker only)."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def _load():
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
    lib = C.CDLL(LIB)
    lib.orc_halfrotate_cholesky.restype = C.c_long
    lib.orc_halfrotate_begin.restype = C.c_void_p
    lib.orc_halfrotate_take.restype = None
    lib.orc_product.restype = None
    return lib


lib = _load()

ORC_NONFINITE = -6  # orc_run status of a non-finite field draw (AFQMC_ERR_NONFINITE)


def set_threads(n):
    """OpenMP threads of the oracle's walker-parallel loops (n <= 0: all)."""
    lib.orc_set_threads(int(n))


def get_threads():
    return lib.orc_get_threads()


def dp(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.view(np.float64).ctypes.data_as(_D) if np.iscomplexobj(a) else a.ctypes.data_as(_D)


def ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


def cz(z):
    return (C.c_double * 2)(complex(z).real, complex(z).imag)


def csrmm(transa, M, N, K, alpha, val, indx, rowptr, B, beta, C_):
    """mySPBLAS::csrmm (sparse.hpp:159-216) on row-major B [.,N] and C [.,N]; C_ updated in place."""
    val = np.ascontiguousarray(val, np.complex128)
    indx = np.ascontiguousarray(indx, np.int32)
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    B = np.ascontiguousarray(B, np.complex128)
    assert C_.dtype == np.complex128 and C_.flags["C_CONTIGUOUS"]
    pe = np.ascontiguousarray(rowptr[1:])
    lib.orc_csrmm(C.c_char(transa.encode()), M, N, K, cz(alpha), dp(val), ip(indx), ip(rowptr), ip(pe), dp(B),
                  B.shape[1], cz(beta), dp(C_), C_.shape[1])
    return C_


def product(opA, opB, A, B, alpha=1.0, beta=0.0, Cm=None):
    A = np.ascontiguousarray(A, np.complex128)
    B = np.ascontiguousarray(B, np.complex128)
    m = A.shape[0] if opA == "N" else A.shape[1]
    k = A.shape[1] if opA == "N" else A.shape[0]
    n = B.shape[1] if opB == "N" else B.shape[0]
    if Cm is None:
        Cm = np.zeros((m, n), np.complex128)
    lib.orc_product(C.c_char(opA.encode()), C.c_char(opB.encode()), m, n, k, cz(alpha), dp(A), A.shape[1], dp(B),
                    B.shape[1], cz(beta), dp(Cm), n)
    return Cm


def getrf(A):
    """ma::getrf on a row-major array: LU of the column-major view (= A^T)."""
    A = np.ascontiguousarray(A, np.complex128).copy()
    piv = np.zeros(min(A.shape), np.int32)
    info = lib.orc_getrf(A.shape[0], A.shape[1], dp(A), A.shape[1], ip(piv))
    return A, piv, info


def invert(A):
    A = np.ascontiguousarray(A, np.complex128).copy()
    piv = np.zeros(A.shape[0], np.int32)
    det = np.zeros(2)
    info = lib.orc_invert(A.shape[0], dp(A), ip(piv), det.ctypes.data_as(_D))
    return A, complex(det[0], det[1]), info


def determinant(A):
    A = np.ascontiguousarray(A, np.complex128).copy()
    piv = np.zeros(A.shape[0], np.int32)
    det = np.zeros(2)
    info = lib.orc_determinant(A.shape[0], dp(A), ip(piv), det.ctypes.data_as(_D))
    return complex(det[0], det[1]), info


def lq(W):
    """ma::gelqf + ma::glq on a row-major W [M][N] (afqmc_sys::orthogonalize)."""
    A = np.ascontiguousarray(W, np.complex128).copy()
    tau = np.zeros(max(A.shape), np.complex128)
    lib.orc_gelqf(A.shape[0], A.shape[1], dp(A), A.shape[1], dp(tau))
    R = A.copy()
    lib.orc_glq(A.shape[0], A.shape[1], dp(A), A.shape[1], dp(tau))
    return A, R, tau


def gelqf(W):
    A = np.ascontiguousarray(W, np.complex128).copy()
    tau = np.zeros(max(A.shape), np.complex128)
    lib.orc_gelqf(A.shape[0], A.shape[1], dp(A), A.shape[1], dp(tau))
    return A, tau


def mixed_density_matrix(conjA, B, compact=True):
    M, N = B.shape
    Cm = np.zeros((N if compact else M, M), np.complex128)
    ov = np.zeros(2)
    info = lib.orc_mixed_density_matrix(M, N, dp(np.ascontiguousarray(conjA)), dp(np.ascontiguousarray(B)), dp(Cm),
                                         int(compact), ov.ctypes.data_as(_D))
    return Cm, complex(ov[0], ov[1]), info


def overlap(conjA, B):
    M, N = B.shape
    ov = np.zeros(2)
    info = lib.orc_overlap(M, N, dp(np.ascontiguousarray(conjA)), dp(np.ascontiguousarray(B)), ov.ctypes.data_as(_D))
    return complex(ov[0], ov[1]), info


def calculate_mixed_density_matrix(sysm, W, Wd, compact=True):
    nw = W.shape[0]
    M, N = sysm.nmo, sysm.naea
    G = np.zeros((2 * (N if compact else M) * M, nw), np.complex128)
    info = lib.orc_calculate_mixed_density_matrix(M, N, nw, dp(sysm.trial_alpha), dp(sysm.trial_beta), dp(W), dp(Wd),
                                                  dp(G), int(compact))
    return G, info


def calculate_overlaps(sysm, W, Wd):
    return lib.orc_calculate_overlaps(sysm.nmo, sysm.naea, W.shape[0], dp(sysm.trial_alpha), dp(sysm.trial_beta),
                                      dp(W), dp(Wd))


def halfrotate(sysm, cutoff=1e-6):
    from miniafqmc_amd.system import Csr
    M, N, K = sysm.nmo, sysm.naea, sysm.nchol
    sp = sysm.spvn
    args = (M, N, K, dp(sysm.trial_alpha), dp(sysm.trial_beta), ip(sp.rowptr), ip(sp.col), dp(sp.val), C.c_double(cutoff))
    nnz = C.c_long(0)
    h = lib.orc_halfrotate_begin(*args, C.byref(nnz))
    rp = np.zeros(K + 1, np.int32)
    col = np.zeros(nnz.value, np.int32)
    val = np.zeros(nnz.value, np.complex128)
    lib.orc_halfrotate_take(C.c_void_p(h), ip(rp), ip(col), dp(val))
    return Csr(K, 2 * N * M, rp, col, val)


def get_vbias(spT, G, transposed=True):
    nw = G.shape[1]
    rows = spT.nrows
    v = np.zeros((rows if transposed else spT.ncols, nw), np.complex128)
    r = lib.orc_get_vbias(spT.nrows, spT.ncols, ip(spT.rowptr), ip(spT.col), dp(spT.val), G.shape[0], nw,
                          dp(np.ascontiguousarray(G)), dp(v), int(transposed))
    assert r == 0
    return v


def get_vhs(sp, X):
    nw = X.shape[1]
    v = np.zeros((sp.nrows, nw), np.complex128)
    lib.orc_get_vhs(sp.nrows, sp.ncols, ip(sp.rowptr), ip(sp.col), dp(sp.val), nw, dp(np.ascontiguousarray(X)), dp(v))
    return v


def apply_expM(V, S, order=6):
    S = np.ascontiguousarray(S, np.complex128).copy()
    lib.orc_apply_expM(S.shape[0], S.shape[1], dp(np.ascontiguousarray(V, np.complex128)), dp(S), order)
    return S


def propagate(sysm, W, vHS):
    W = np.ascontiguousarray(W, np.complex128).copy()
    lib.orc_propagate(sysm.nmo, sysm.naea, W.shape[0], dp(sysm.propg1), dp(np.ascontiguousarray(vHS)), dp(W))
    return W


def orthogonalize(W):
    W = np.ascontiguousarray(W, np.complex128).copy()
    nw, _, M, N = W.shape
    lib.orc_orthogonalize(M, N, nw, dp(W))
    return W


def calculate_energy(sysm, Wd, G):
    nw = G.shape[1]
    e = C.c_double(0)
    v = sysm.vakbl
    lib.orc_calculate_energy(sysm.nmo, sysm.naea, nw, dp(Wd), dp(np.ascontiguousarray(G)), dp(sysm.haj), v.nrows,
                             ip(v.rowptr), ip(v.col), dp(v.val), C.byref(e))
    return e.value


def dense_Lt(sysm):
    """Unscaled half-rotated Cholesky matrix, dense [K][2NM], no cutoff."""
    spT = halfrotate(sysm, cutoff=-1.0)
    return np.ascontiguousarray(spT.dense() / np.sqrt(sysm.dt))


def energy_cholesky(sysm, Wd, G, Lt=None):
    if Lt is None:
        Lt = dense_Lt(sysm)
    nw = G.shape[1]
    e = C.c_double(0)
    lib.orc_energy_cholesky(sysm.nmo, sysm.naea, sysm.nchol, nw, dp(Wd), dp(np.ascontiguousarray(G)), dp(sysm.haj),
                            dp(np.ascontiguousarray(Lt)), C.byref(e))
    return e.value


def fields(vbias, x):
    K, nw = vbias.shape
    X = np.zeros((K, nw), np.complex128)
    h = np.zeros(nw, np.complex128)
    lib.orc_fields(K, nw, dp(np.ascontiguousarray(vbias)), dp(np.ascontiguousarray(x, np.float64)), dp(X), dp(h))
    return X, h


def weights(dt, hybridW, Wd, eshift):
    e = C.c_double(eshift)
    lib.orc_weights(Wd.shape[0], C.c_double(dt), dp(np.ascontiguousarray(hybridW)), dp(Wd), C.byref(e))
    return e.value


def branch(W, Wd, u):
    """Comb population control in place; returns (status, parent)."""
    nw, _, M, N = W.shape
    parent = np.zeros(nw, np.int32)
    st = lib.orc_branch(M, N, nw, C.c_double(u), dp(W), dp(Wd), ip(parent))
    return st, parent


def compress(nrows, rows, cols, vals):
    rows = np.ascontiguousarray(rows, np.int32).copy()
    cols = np.ascontiguousarray(cols, np.int32).copy()
    vals = np.ascontiguousarray(vals, np.complex128).copy()
    rp = np.zeros(nrows + 1, np.int32)
    lib.orc_compress(nrows, C.c_long(len(rows)), ip(rows), ip(cols), dp(vals), ip(rp))
    return rows, cols, vals, rp


def run(sysm, nwalk, nsteps, nsubsteps, northo, x_all, energy_mode=None, spT=None, Lt=None, cutoff=1e-6,
        transposed=True):
    """Whole driver loop (miniafqmc_base.cpp:213-349).  Returns (eav[nsteps+1], W, Wd, eshift, timers).
    transposed=False takes the driver's non-transposed branch (full G, T(Spvn))."""
    M, N, K = sysm.nmo, sysm.naea, sysm.nchol
    if spT is None and transposed:
        spT = halfrotate(sysm, cutoff)
    if energy_mode is None:
        energy_mode = 0 if sysm.vakbl is not None else 1
    if energy_mode == 1 and Lt is None:
        Lt = dense_Lt(sysm)
    W = np.zeros((nwalk, 2, M, N), np.complex128)
    Wd = np.zeros((nwalk, 8), np.complex128)
    eav = np.zeros(nsteps + 1)
    es = C.c_double(0)
    tm = np.zeros(11)
    sp = sysm.spvn
    v = sysm.vakbl
    x_all = np.ascontiguousarray(x_all, np.float64)
    assert x_all.shape == (nsteps * nsubsteps, K, nwalk)
    dummy_i = np.zeros(1, np.int32)
    dummy_z = np.zeros(1, np.complex128)
    info = lib.orc_run(
        M, N, K, nwalk, nsteps, nsubsteps, northo, C.c_double(sysm.dt), dp(sysm.trial_alpha), dp(sysm.trial_beta),
        dp(sysm.propg1), dp(sysm.haj), ip(sp.rowptr), ip(sp.col), dp(sp.val),
        ip(spT.rowptr) if transposed else None, ip(spT.col) if transposed else None,
        dp(spT.val) if transposed else None, energy_mode, ip(v.rowptr if v is not None else dummy_i), ip(v.col if v is not None else dummy_i),
        dp(v.val if v is not None else dummy_z), dp(Lt if Lt is not None else dummy_z), dp(x_all), dp(W), dp(Wd),
        eav.ctypes.data_as(_D), C.byref(es), tm.ctypes.data_as(_D))
    return eav, W, Wd, es.value, tm, info
