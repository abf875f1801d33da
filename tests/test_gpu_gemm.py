"""GPU complex GEMM (FP64 / FP32 / split-BF16 MFMA) against numpy.

Checks each precision's lane layouts on every tile candidate and operand
orientation (with the conjugated B forms), including ragged edges.  Tolerances are relative to
max|C|: FP64 1e-12, FP32 2e-5, split-BF16 5e-5 (hi+lo keeps ~16 mantissa
bits per operand; products accumulate in FP32).
"""
import ctypes as C

import numpy as np
import pytest

from miniafqmc_amd import _native as nat

pytestmark = pytest.mark.gpu
TOL = {0: 1e-12, 1: 2e-5, 2: 5e-5}
_D = C.POINTER(C.c_double)


def zgemm(prec, opA, opB, A, B, alpha, beta, Cm):
    m, n = Cm.shape
    k = A.shape[1] if opA == "N" else A.shape[0]
    a = np.array([alpha.real, alpha.imag])
    b = np.array([beta.real, beta.imag])
    out = Cm.copy()
    r = nat.lib.afqmc_zgemm(prec, opA.encode(), opB.encode(), m, n, k, a.ctypes.data_as(_D),
                            A.view(np.float64).ctypes.data_as(_D), A.shape[1], B.view(np.float64).ctypes.data_as(_D),
                            B.shape[1], b.ctypes.data_as(_D), out.view(np.float64).ctypes.data_as(_D), n)
    assert r == 0
    return out


@pytest.mark.parametrize("prec", [0, 1, 2, 0 | 1 << 8, 0 | 2 << 8, 0 | 3 << 8, 1 | 1 << 8, 1 | 2 << 8, 1 | 3 << 8,
                                  2 | 1 << 8, 2 | 2 << 8, 2 | 3 << 8, 2 | 4 << 8, 2 | 5 << 8])
@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T"), ("N", "C"), ("T", "R")])
@pytest.mark.parametrize("m,n,k", [(16, 16, 64), (7, 5, 3), (33, 17, 70), (130, 70, 40), (256, 128, 96)])
def test_zgemm(require_gpu, prec, opA, opB, m, n, k):
    r = np.random.default_rng(m * 1000 + n * 10 + k)
    A = r.standard_normal((m, k) if opA == "N" else (k, m)) + 1j * r.standard_normal((m, k) if opA == "N" else (k, m))
    bshape = (k, n) if opB in ("N", "R") else (n, k)
    B = r.standard_normal(bshape) + 1j * r.standard_normal(bshape)
    C0 = r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))
    alpha, beta = 0.7 - 0.3j, -0.5 + 1.25j
    opa = A if opA == "N" else A.T
    opb = {"N": B, "T": B.T, "C": B.conj().T, "R": B.conj()}[opB]
    ref = alpha * opa @ opb + beta * C0
    out = zgemm(prec, opA, opB, np.ascontiguousarray(A), np.ascontiguousarray(B), alpha, beta, C0)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    assert err < TOL[prec & 0xff], f"prec {prec}: rel err {err:.2e}"
