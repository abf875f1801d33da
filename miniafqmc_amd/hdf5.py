"""The miniapp's HDF5 input (afqmc.h5) in the reference's format, through
libafqmc_h5.so (host code over the HDF5 C library; built only where HDF5 is
installed, and not needed by the GPU path).

``read_hdf5(path, dt)`` returns what ``afqmc::Initialize``
(src/Matrix/initialize_serial.hpp:35-228) hands to the driver, as a
``System``; ``write_hdf5(system, path)`` writes a system in the same format
(the file holds the Cholesky vectors without the sqrt(dt) factor).
"""
import ctypes as C
import math
import os

import numpy as np

from . import _native as nat
from .system import Csr, System, _arr, _csr

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libafqmc_h5.so")
_lib = None


def available():
    """True when libafqmc_h5.so is built and loads (the HDF5 C library is present)."""
    try:
        _load()
        return True
    except OSError:
        return False


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise OSError(f"{_PATH} is not built (it needs the HDF5 C library at build time)")
        lib = C.CDLL(_PATH)
        lib.afqmc_h5_read.argtypes = [C.c_char_p, C.c_double, C.POINTER(C.c_void_p)]
        lib.afqmc_h5_get.argtypes = [C.c_void_p, C.POINTER(nat.SynthView)]
        lib.afqmc_h5_destroy.argtypes = [C.c_void_p]
        lib.afqmc_h5_write.argtypes = [C.c_char_p, C.POINTER(nat.SynthView), C.c_long]
        lib.afqmc_h5_error.restype = C.c_char_p
        _lib = lib
    return _lib


def _check(lib, code):
    if code != 0:
        raise nat.AfqmcError(code, lib.afqmc_h5_error().decode())


def read_hdf5(path, dt) -> System:
    """afqmc::Initialize (initialize_serial.hpp:35-228): conj trial, compacted
    haj and Vakbl, Propg1 and the sqrt(dt)-scaled Spvn, sorted CSR."""
    lib = _load()
    h = C.c_void_p()
    _check(lib, lib.afqmc_h5_read(os.fsencode(path), float(dt), C.byref(h)))
    try:
        v = nat.SynthView()
        _check(lib, lib.afqmc_h5_get(h, C.byref(v)))
        M, N, K = v.nmo, v.naea, v.nchol
        vak = _csr(v.vakbl) if v.vakbl.nnz > 0 else None
        return System(M, N, K, float(dt), 0,
                      _arr(v.trial_alpha, M * N).reshape(M, N),
                      _arr(v.trial_beta, M * N).reshape(M, N),
                      _arr(v.haj, 2 * N * M).reshape(2 * N, M),
                      _arr(v.propg1, M * M).reshape(M, M),
                      None, _csr(v.spvn), vak, dict(file=os.fspath(path)))
    finally:
        lib.afqmc_h5_destroy(h)


def write_hdf5(system: System, path, spvn_block=0):
    """Write `system` (whose Spvn carries sqrt(dt)) in the reference's format;
    spvn_block: entries per Spvn_index_<b>/Spvn_vals_<b> block (0: one)."""
    lib = _load()
    s = system
    raw = Csr(s.spvn.nrows, s.spvn.ncols, np.ascontiguousarray(s.spvn.rowptr, np.int32),
              np.ascontiguousarray(s.spvn.col, np.int32), s.spvn.val / math.sqrt(s.dt))
    keep = [np.ascontiguousarray(a, np.complex128) for a in (s.trial_alpha, s.trial_beta, s.haj, s.propg1)]
    dp = [a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)) for a in keep]
    v = nat.SynthView()
    v.nmo, v.naea, v.nchol = s.nmo, s.naea, s.nchol
    v.trial_alpha, v.trial_beta, v.haj, v.propg1 = dp
    v.spvn = raw.view()
    if s.vakbl is not None and s.vakbl.nnz > 0:
        vk = Csr(s.vakbl.nrows, s.vakbl.ncols, np.ascontiguousarray(s.vakbl.rowptr, np.int32),
                 np.ascontiguousarray(s.vakbl.col, np.int32), np.ascontiguousarray(s.vakbl.val))
        v.vakbl = vk.view()
    _check(lib, lib.afqmc_h5_write(os.fsencode(path), C.byref(v), int(spvn_block)))
