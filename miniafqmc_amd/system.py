"""Synthetic AFQMC systems (the reference's example afqmc.h5 is not shipped).

``make_system(...)`` returns numpy copies of what ``afqmc::Initialize``
(src/Matrix/initialize_serial.hpp:35-228) hands to the driver: conj trial
orbitals per spin, compacted ``haj``, ``Propg1``, the sqrt(dt)-scaled CSR
``Spvn`` and optionally the compacted CSR ``Vakbl``.  ``fields(...)`` makes
the real N(0,1) auxiliary-field draws of one substep, ``[K][nwalk]``.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _native as nat


@dataclass
class Csr:
    nrows: int
    ncols: int
    rowptr: np.ndarray  # int32 [nrows+1]
    col: np.ndarray     # int32 [nnz]
    val: np.ndarray     # complex128 [nnz]

    @property
    def nnz(self):
        return int(self.rowptr[-1] - self.rowptr[0])

    def view(self):
        v = nat.CsrView()
        v.nrows, v.ncols, v.nnz = self.nrows, self.ncols, self.nnz
        v.rowptr = self.rowptr.ctypes.data_as(C.POINTER(C.c_int))
        v.col = self.col.ctypes.data_as(C.POINTER(C.c_int))
        v.val = self.val.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        return v

    def dense(self):
        out = np.zeros((self.nrows, self.ncols), np.complex128)
        rows = np.repeat(np.arange(self.nrows), np.diff(self.rowptr))
        np.add.at(out, (rows, self.col), self.val)  # repeated entries add up, as in csrmm
        return out


@dataclass
class System:
    nmo: int
    naea: int
    nchol: int
    dt: float
    seed: int
    trial_alpha: np.ndarray  # [M][N] conj(Psi_a)
    trial_beta: np.ndarray
    haj: np.ndarray          # [2N][M]
    propg1: np.ndarray       # [M][M]
    h1: np.ndarray           # [M][M]
    spvn: Csr                # [M*M][K], sqrt(dt) scaled
    vakbl: Optional[Csr] = None
    meta: dict = field(default_factory=dict)


def _arr(ptr, n, dtype=np.complex128):
    if n == 0:
        return np.zeros(0, dtype)
    if dtype == np.complex128:
        a = np.ctypeslib.as_array(ptr, shape=(2 * n,)).copy()
        return a.view(np.complex128)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


def _csr(v):
    if v.nnz == 0 and not v.rowptr:
        return None
    rp = _arr(v.rowptr, v.nrows + 1, np.int32)
    return Csr(v.nrows, v.ncols, rp, _arr(v.col, v.nnz, np.int32), _arr(v.val, v.nnz))


def make_system(nmo, naea, nchol, density=1.0, sigma_L=0.1, sigma_h=1.0, dt=0.005, seed=1,
                with_vakbl=False) -> System:
    p = nat.SynthParams(nmo, naea, nchol, density, sigma_L, sigma_h, dt, seed, int(with_vakbl))
    h = C.c_void_p()
    nat.check(nat.lib.afqmc_synth_create(C.byref(p), C.byref(h)))
    try:
        v = nat.SynthView()
        nat.check(nat.lib.afqmc_synth_get(h, C.byref(v)))
        M, N = nmo, naea
        sysm = System(
            nmo, naea, nchol, dt, seed,
            _arr(v.trial_alpha, M * N).reshape(M, N),
            _arr(v.trial_beta, M * N).reshape(M, N),
            _arr(v.haj, 2 * N * M).reshape(2 * N, M),
            _arr(v.propg1, M * M).reshape(M, M),
            _arr(v.h1, M * M).reshape(M, M),
            _csr(v.spvn),
            _csr(v.vakbl) if v.vakbl.nnz > 0 else None,
            dict(density=density, sigma_L=sigma_L, sigma_h=sigma_h),
        )
    finally:
        nat.lib.afqmc_synth_destroy(h)
    return sysm


def fields(seed, nchol, nwalk, substep, w0=0, wtot=None):
    """Real N(0,1) draws [K][nwalk] for walkers w0.. of a run with wtot walkers."""
    wtot = nwalk + w0 if wtot is None else wtot
    x = np.empty((nchol, nwalk), np.float64)
    nat.check(nat.lib.afqmc_synth_fields(seed, nchol, nwalk, w0, wtot, substep,
                                          x.ctypes.data_as(C.POINTER(C.c_double))))
    return x


def fields_block(seed, nchol, nwalk, first, count, w0=0, wtot=None):
    """Draws of substeps first..first+count-1 as [count][K][nwalk]."""
    out = np.empty((count, nchol, nwalk), np.float64)
    for s in range(count):
        out[s] = fields(seed, nchol, nwalk, first + s, w0, wtot)
    return out


# Workload presets, one per configuration of the benchmark definition.
# sigma_L of the larger systems is scaled so that M*K*E|L|^2 (which sets the
# size of vHS) equals the 64-orbital case: E|L|^2 = 2 sigma^2 rho (1 - ln rho).
def _sigma_for(M, K, rho, ref=(64, 256, 0.3, 0.02)):
    import math
    M0, K0, r0, s0 = ref
    e0 = 2 * s0 * s0 * r0 * (1 - math.log(r0))
    e = M0 * K0 * e0 / (M * K)
    return math.sqrt(e / (2 * rho * (1 - math.log(rho))))


PRESETS = {
    # 64 orbitals, 16+16 electrons, 256 vectors at 30% density, 16 walkers
    "N16_M64": dict(nmo=64, naea=16, nchol=256, density=0.3, sigma_L=0.02, nwalk=16, dt=0.005,
                    nsteps=20, nsubsteps=1, northo=5, with_vakbl=True, seed=20261202 + 1000 * 1),
    # 256 orbitals, 64+64 electrons, 1024 vectors at 20% density, 256 walkers
    "N64_M256": dict(nmo=256, naea=64, nchol=1024, density=0.2, sigma_L=_sigma_for(256, 1024, 0.2),
                     nwalk=256, dt=0.005, nsteps=5, nsubsteps=10, northo=10, with_vakbl=False,
                     seed=20261202 + 1000 * 2),
    # 512 orbitals, 128+128 electrons, 2048 dense vectors, 1024 walkers
    "N128_M512": dict(nmo=512, naea=128, nchol=2048, density=1.0, sigma_L=_sigma_for(512, 2048, 1.0),
                      nwalk=1024, dt=0.005, nsteps=1, nsubsteps=100, northo=10, with_vakbl=False,
                      seed=20261202 + 1000 * 4),
}


def preset_system(name, **over):
    p = dict(PRESETS[name])
    p.update(over)
    return make_system(p["nmo"], p["naea"], p["nchol"], p["density"], p["sigma_L"], 1.0, p["dt"],
                       p["seed"], p["with_vakbl"]), p


# The benchmark definition's uneven batch of independent systems:
# M = 96 ... 640, N = M/4, K = 4M, 15-40 % density, 64-512 walkers.
# sigma_L keeps M*K*E|L|^2 at the 64-orbital preset's value.
MULTI = [(96, 64, 0.15), (128, 96, 0.2), (192, 128, 0.25), (256, 160, 0.3), (320, 192, 0.35), (384, 256, 0.4),
         (512, 512, 0.4), (640, 128, 0.4)]


def multi_system(i, dt=0.005):
    """(System, walkers, seed) of member i of the uneven batch."""
    M, nw, rho = MULTI[i]
    N, K = M // 4, 4 * M
    seed = 20261202 + 5000 + i
    s = make_system(M, N, K, density=rho, sigma_L=_sigma_for(M, K, rho), sigma_h=1.0, dt=dt, seed=seed)
    return s, nw, seed


def multi_systems(dt=0.005):
    """[(System, walkers, seed)] of the uneven batch."""
    return [multi_system(i, dt) for i in range(len(MULTI))]
