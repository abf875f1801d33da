"""The allocation-free mixed density matrix (afqmc_mixed_density_matrix_noalloc,
no counterpart in the reference) against the context entry point
afqmc_mixed_density_matrix (afqmc_sys.hpp:95-116), on device buffers the
test allocates through the HIP runtime: G and the overlaps at 1e-12
relative (FP64, the same kernels), on two calls in a row with the second
one smaller, so a stale scratch array from the first call would show.
Requests beyond the scratch array are refused before any launch.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import miniafqmc_amd as m
from miniafqmc_amd import _native as nat

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


class Hip:
    """The few HIP runtime calls the test needs (the caller's allocations)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipStreamCreate",
                     "hipStreamDestroy", "hipStreamSynchronize"):
            getattr(self.lib, name).restype = C.c_int
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipStreamDestroy.argtypes = [C.c_void_p]
        self.lib.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.ptrs.append(p.value)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host -> device
        return p

    def download(self, p, shape, dtype=np.complex128):
        out = np.empty(shape, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device -> host
        return out

    def stream(self):
        s = C.c_void_p()
        assert self.lib.hipStreamCreate(C.byref(s)) == 0
        return s.value

    def free_all(self):
        for p in self.ptrs:
            self.lib.hipFree(p)
        self.ptrs = []


def walkers(s, nw, seed):
    r = np.random.default_rng(seed)
    W = np.zeros((nw, 2, s.nmo, s.naea), complex)
    W[:, 0] = s.trial_alpha.conj()
    W[:, 1] = s.trial_beta.conj()
    W += 0.2 * (r.standard_normal(W.shape) + 1j * r.standard_normal(W.shape))
    return W


def context_result(s, W):
    nw = W.shape[0]
    g = m.AfqmcSys(s, nw, "fp64_csr")
    Wd = np.zeros((nw, 8), complex)
    Wd[:, 1] = 1.0
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, Wd)
    g.calculate_mixed_density_matrix()
    G, ov = g.get(nat.BUF_G), g.W_data[:, 2:4].copy()
    g.close()
    return G, ov


@pytest.mark.parametrize("shape", [(64, 16, 32, (48, 20)), (512, 128, 16, (128, 37))])
def test_noalloc_matches_context_entry_point(require_gpu, shape):
    M, N, K, sizes = shape
    s = m.make_system(M, N, K, density=1.0, sigma_L=0.01, sigma_h=1.0, dt=0.005, seed=4242 + M)
    hip = Hip()
    try:
        trial = np.stack([s.trial_alpha, s.trial_beta])  # trialwfn = conj(Psi) per spin
        d_trial = hip.upload(trial)
        stream = hip.stream()
        for call, nw in enumerate(sizes):
            say(f"no-alloc mixed density matrix, M={M} N={N}, call {call + 1}: {nw} walkers")
            W = walkers(s, nw, seed=100 * call + M)
            Gref, ovref = context_result(s, W)
            d_W = hip.upload(W)
            d_G = hip.alloc(2 * N * M * nw * 16)
            d_ov = hip.alloc(2 * nw * 16)
            st = stream if call == 1 else None  # the null stream, then a created one
            nat.check(nat.lib.afqmc_mixed_density_matrix_noalloc(M, N, nw, d_trial, d_W, d_G, d_ov, st))
            assert hip.lib.hipDeviceSynchronize() == 0
            G = hip.download(d_G, (2 * N * M, nw))
            ov = hip.download(d_ov, (nw, 2))
            eg, eo = rel(G, Gref), rel(ov, ovref)
            say(f"  G rel err {eg:.2e}, overlaps rel err {eo:.2e}")
            assert eg < 1e-12 and eo < 1e-12
        hip.lib.hipStreamDestroy(stream)
    finally:
        hip.free_all()


def test_noalloc_rejects_oversized_requests(require_gpu):
    hip = Hip()
    try:
        M, N = 512, 128
        per_walker = 2 * N * N * 16
        nw = nat.NOALLOC_SCRATCH_BYTES // per_walker + 1
        say(f"no-alloc request beyond the scratch array: {nw} walkers at N={N}")
        d = hip.alloc(16)  # the call is refused before any pointer is used
        assert nat.lib.afqmc_mixed_density_matrix_noalloc(M, N, nw, d, d, d, d, None) == 2  # AFQMC_ERR_SHAPE
        assert nat.lib.afqmc_mixed_density_matrix_noalloc(M, 129, 1, d, d, d, d, None) == 8  # AFQMC_ERR_UNSUPPORTED
        assert nat.lib.afqmc_mixed_density_matrix_noalloc(M, N, 0, d, d, d, d, None) == 1  # AFQMC_ERR_ARG
    finally:
        hip.free_all()
