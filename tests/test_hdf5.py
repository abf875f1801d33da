"""HDF5 input in the reference's afqmc.h5 format (libafqmc_h5.so; CPU only).

- A file crafted dataset by dataset through the HDF5 C library (not through
  this build's writer) is read with the semantics of afqmc::Initialize
  (src/Matrix/initialize_serial.hpp:35-228): one-body rows compacted with
  a = i (alpha) or i - M + NAEA (beta), trial = conj(Wavefun), Vakbl rows and
  columns compacted, Spvn gathered over blocks, sorted by (row, column) with
  duplicates kept, and scaled by sqrt(dt).  The expected arrays are written
  out by hand.
- Malformed files fail where the reference asserts.
- write -> read reproduces a synthetic system, and the oracle gives the same
  walker step on the system read back.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import miniafqmc_amd as m
from miniafqmc_amd import _native as nat
from miniafqmc_amd import hdf5

pytestmark = pytest.mark.skipif(not hdf5.available(), reason="libafqmc_h5.so not built (no HDF5 C library)")


class H5Craft:
    """Minimal writer over the HDF5 C library already loaded by libafqmc_h5.so."""

    def __init__(self, path):
        hdf5._load()
        so = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libhdf5.so" in line or "libhdf5-" in line:
                    so = line.split()[-1]
                    break
        assert so, "libhdf5 not mapped"
        L = self.L = C.CDLL(so)
        L.H5open()
        for fn in ("H5Fcreate", "H5Gcreate2", "H5Screate_simple", "H5Dcreate2"):
            getattr(L, fn).restype = C.c_int64
        L.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, C.c_int64, C.c_int64]
        L.H5Gcreate2.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_int64, C.c_int64]
        L.H5Screate_simple.argtypes = [C.c_int, C.POINTER(C.c_uint64), C.c_void_p]
        L.H5Dcreate2.argtypes = [C.c_int64, C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64]
        L.H5Dwrite.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
        for fn in ("H5Dclose", "H5Sclose", "H5Gclose", "H5Fclose"):
            getattr(L, fn).argtypes = [C.c_int64]
        self.t = {k: C.c_int64.in_dll(L, f"H5T_NATIVE_{k}_g").value for k in ("INT", "LONG", "DOUBLE")}
        self.f = L.H5Fcreate(os.fsencode(path), 2, 0, 0)  # H5F_ACC_TRUNC
        assert self.f >= 0
        self.groups = {}

    def group(self, path):
        parent = self.f
        for part in path.split("/"):
            key = (parent, part)
            if key not in self.groups:
                self.groups[key] = self.L.H5Gcreate2(parent, part.encode(), 0, 0, 0)
            parent = self.groups[key]
        return parent

    def put(self, grp, name, arr):
        arr = np.ascontiguousarray(arr)
        if arr.dtype == np.complex128:
            arr = arr.view(np.float64).reshape(-1, 2)
        tkey = {np.dtype(np.int32): "INT", np.dtype(np.int64): "LONG", np.dtype(np.float64): "DOUBLE"}[arr.dtype]
        dims = (C.c_uint64 * arr.ndim)(*arr.shape)
        sp = self.L.H5Screate_simple(arr.ndim, dims, None)
        d = self.L.H5Dcreate2(grp, name.encode(), self.t[tkey], sp, 0, 0, 0)
        assert d >= 0, name
        if arr.size:
            assert self.L.H5Dwrite(d, self.t[tkey], 0, 0, 0, arr.ctypes.data) >= 0
        self.L.H5Dclose(d)
        self.L.H5Sclose(sp)

    def close(self):
        for g in reversed(list(self.groups.values())):
            self.L.H5Gclose(g)
        self.L.H5Fclose(self.f)


# A hand-made 3-orbital, 1+1-electron, 2-vector file.
M, N, K = 3, 1, 2
HIJ_INDX = np.array([0 * 3 + 1, 3 * 3 + 2, 0 * 3 + 0], np.int32)  # alpha (0,1), beta i=M (row a=1) (.,2), alpha (0,0)
HIJ = np.array([1 + 2j, 3 - 1j, -0.5 + 0j])
WAVEFUN = np.array([0.1 + 0.2j, 0.3 - 0.4j, 0.5 + 0j, -0.6 + 0.1j, 0.7 + 0.7j, 0.2 - 0.3j])
# two-body entries at expanded (row, col): (10 = beta i=3, j=1 ; 2 = alpha i=0, j=2) and (0 ; 9 = beta i=3, j=0)
SPH_ROWINDEX = np.zeros(2 * M * M + 1, np.int32)
SPH_ROWINDEX[1:] = 1      # row 0 holds one entry
SPH_ROWINDEX[11:] = 2     # row 10 holds the second
SPH_COLS = np.array([9, 2], np.int32)
SPH_VALS = np.array([0.25 - 0.5j, -1.5 + 0.75j])
PROPG1 = (np.arange(9) * (1 + 0.5j)).astype(np.complex128)
# Spvn over two blocks, unsorted, with a repeated (row, col)
BLK0_IDX, BLK0_VAL = np.array([5, 1, 2, 0], np.int32), np.array([1 + 1j, 2 + 0j])
BLK1_IDX, BLK1_VAL = np.array([2, 0, 0, 1], np.int32), np.array([3 - 1j, 4 + 0.5j])
DT = 0.01


def craft(path, dims=None, hij_indx=HIJ_INDX, blk1_idx=BLK1_IDX, skip_propagators=False):
    w = H5Craft(path)
    g = w.group("Wavefunctions/PureSingleDeterminant")
    d = np.array([0, 2, 2 * M * M, 2 * M * M, M, N, N, 0], np.int32) if dims is None else np.asarray(dims, np.int32)
    w.put(g, "dims", d)
    w.put(g, "hij_indx", hij_indx)
    w.put(g, "hij", HIJ)
    w.put(g, "Wavefun", WAVEFUN)
    w.put(g, "SpHijkl_vals", SPH_VALS)
    w.put(g, "SpHijkl_cols", SPH_COLS)
    w.put(g, "SpHijkl_rowIndex", SPH_ROWINDEX)
    if not skip_propagators:
        p = w.group("Propagators/phaseless_ImpSamp_ForceBias")
        w.put(p, "Spvn_dims", np.array([4, M * M, K, M, 2], np.int64))
        w.put(p, "Spvn_propg1", PROPG1)
        w.put(p, "Spvn_block_sizes", np.array([2, 2], np.int32))
        w.put(p, "Spvn_index_0", BLK0_IDX)
        w.put(p, "Spvn_vals_0", BLK0_VAL)
        w.put(p, "Spvn_index_1", blk1_idx)
        w.put(p, "Spvn_vals_1", BLK1_VAL)
    w.close()


def test_reader_semantics_on_crafted_file(tmp_path):
    path = tmp_path / "crafted.h5"
    craft(path)
    s = hdf5.read_hdf5(path, DT)
    assert (s.nmo, s.naea, s.nchol) == (M, N, K)
    # trial = conj(Wavefun) per spin, [M][N]
    np.testing.assert_array_equal(s.trial_alpha, WAVEFUN[:3].conj().reshape(3, 1))
    np.testing.assert_array_equal(s.trial_beta, WAVEFUN[3:].conj().reshape(3, 1))
    # haj [2N][M]: alpha row 0 and beta row (i = M) -> a = 1
    haj = np.zeros((2, 3), complex)
    haj[0, 1], haj[1, 2], haj[0, 0] = HIJ
    np.testing.assert_array_equal(s.haj, haj)
    np.testing.assert_array_equal(s.propg1, PROPG1.reshape(3, 3))
    # Vakbl compacted: expanded 0 -> 0, 9 (i=3, j=0) -> a=1 -> 3, 10 (i=3, j=1) -> 4, 2 -> 2
    V = np.zeros((6, 6), complex)
    V[0, 3] = SPH_VALS[0]
    V[4, 2] = SPH_VALS[1]
    assert (s.vakbl.nrows, s.vakbl.ncols) == (6, 6)
    np.testing.assert_array_equal(s.vakbl.dense(), V)
    # Spvn: gathered over blocks, sorted by (row, col), duplicates kept in order, times sqrt(dt)
    q = math.sqrt(DT)
    np.testing.assert_array_equal(s.spvn.rowptr, [0, 1, 1, 3, 3, 3, 4, 4, 4, 4])
    np.testing.assert_array_equal(s.spvn.col, [1, 0, 0, 1])
    np.testing.assert_array_equal(s.spvn.val, np.array([4 + 0.5j, 2 + 0j, 3 - 1j, 1 + 1j]) * q)
    assert s.h1 is None


@pytest.mark.parametrize("case", ["naeb", "unrestricted", "row", "spvn_index", "group", "dt", "not_hdf5"])
def test_reader_rejects_malformed_files(tmp_path, case):
    path = tmp_path / f"{case}.h5"
    dt = DT
    if case == "naeb":
        craft(path, dims=[0, 2, 18, 18, M, N, N + 1, 0])
    elif case == "unrestricted":
        craft(path, dims=[0, 2, 18, 18, M, N, N, 1])
    elif case == "row":  # i = 1 is not an occupied row for N = 1
        craft(path, hij_indx=np.array([1 * 3 + 0, 0, 1], np.int32))
    elif case == "spvn_index":  # column K is out of range
        craft(path, blk1_idx=np.array([2, 0, 0, K], np.int32))
    elif case == "group":
        craft(path, skip_propagators=True)
    elif case == "dt":
        craft(path)
        dt = 0.0
    else:
        path.write_text("not an hdf5 file\n")
    with pytest.raises(nat.AfqmcError) as ei:
        hdf5.read_hdf5(path, dt)
    expect = {"naeb": 2, "unrestricted": 2, "row": 2, "spvn_index": 2, "group": 1, "dt": 1, "not_hdf5": 1}[case]
    assert ei.value.code == expect, str(ei.value)


@pytest.fixture(scope="module")
def sysm():
    return m.make_system(64, 16, 256, density=0.3, sigma_L=0.02, sigma_h=1.0, dt=0.005, seed=7, with_vakbl=True)


@pytest.mark.parametrize("block", [0, 1000])
def test_write_read_round_trip(tmp_path, sysm, block):
    s = sysm
    path = tmp_path / "sys.h5"
    hdf5.write_hdf5(s, path, spvn_block=block)
    r = hdf5.read_hdf5(path, s.dt)
    assert (r.nmo, r.naea, r.nchol) == (s.nmo, s.naea, s.nchol)
    for name in ("trial_alpha", "trial_beta", "haj", "propg1"):
        np.testing.assert_array_equal(getattr(r, name), getattr(s, name), err_msg=name)
    np.testing.assert_array_equal(r.spvn.rowptr, s.spvn.rowptr)
    np.testing.assert_array_equal(r.spvn.col, s.spvn.col)
    # the file holds L = Spvn / sqrt(dt); reading multiplies back
    np.testing.assert_allclose(r.spvn.val, s.spvn.val, rtol=4e-16, atol=0)
    np.testing.assert_array_equal(r.vakbl.rowptr, s.vakbl.rowptr)
    np.testing.assert_array_equal(r.vakbl.col, s.vakbl.col)
    np.testing.assert_array_equal(r.vakbl.val, s.vakbl.val)


def test_oracle_step_same_on_read_system(tmp_path, sysm):
    """The oracle's walker step and energies on the system read back equal
    those on the original (to the sqrt(dt) round trip of Spvn)."""
    import oracle_lib as orc
    s = sysm
    path = tmp_path / "sys.h5"
    hdf5.write_hdf5(s, path)
    r = hdf5.read_hdf5(path, s.dt)
    nw = 4
    x = m.fields_block(11, s.nchol, nw, 0, 3)
    e0 = orc.run(s, nw, 1, 3, 2, x)[0]
    e1 = orc.run(r, nw, 1, 3, 2, x)[0]
    np.testing.assert_allclose(e1, e0, rtol=1e-12)


H5DUMP = next((p for p in ("/opt/conda/bin/h5dump", "/usr/bin/h5dump") if os.path.exists(p)), None)


def _h5dump_data(path, dset):
    import re
    import subprocess
    out = subprocess.run([H5DUMP, "-y", "-w", "0", "-d", dset, str(path)], capture_output=True, text=True,
                         check=True).stdout
    body = out[out.index("DATA {") + len("DATA {"):]
    body = body[:body.index("}")]
    return np.array([float(t) for t in re.findall(r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", body)])


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_writer_layout_via_h5dump(tmp_path):
    """The written file read by HDF5's own h5dump: dataset types and shapes as
    the reference's reader takes them (int, long, complex as double [n][2]) and
    the one-body rows expanded back to spin-orbital indices."""
    import subprocess
    s = m.make_system(8, 2, 6, density=0.5, sigma_L=0.05, sigma_h=1.0, dt=0.005, seed=3, with_vakbl=True)
    path = tmp_path / "t.h5"
    hdf5.write_hdf5(s, path, spvn_block=100)
    hdr = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, check=True).stdout
    nnz, nv = s.spvn.nnz, s.vakbl.nnz
    expect = {
        "dims": ("H5T_STD_I32LE", "( 8 )"), "hij_indx": ("H5T_STD_I32LE", "( 32 )"),
        "hij": ("H5T_IEEE_F64LE", "( 32, 2 )"), "Wavefun": ("H5T_IEEE_F64LE", "( 32, 2 )"),
        "SpHijkl_vals": ("H5T_IEEE_F64LE", f"( {nv}, 2 )"), "SpHijkl_cols": ("H5T_STD_I32LE", f"( {nv} )"),
        "SpHijkl_rowIndex": ("H5T_STD_I32LE", "( 129 )"), "Spvn_dims": ("H5T_STD_I64LE", "( 5 )"),
        "Spvn_propg1": ("H5T_IEEE_F64LE", "( 64, 2 )"), "Spvn_index_0": ("H5T_STD_I32LE", "( 200 )"),
        "Spvn_vals_0": ("H5T_IEEE_F64LE", "( 100, 2 )"),
    }
    for name, (dtype, shape) in expect.items():
        blk = hdr[hdr.index(f'DATASET "{name}"'):]
        blk = blk[:blk.index("}")]
        assert dtype in blk and f"SIMPLE {{ {shape} /" in blk, (name, blk)
    g = "/Wavefunctions/PureSingleDeterminant/"
    np.testing.assert_array_equal(_h5dump_data(path, g + "dims"), [0, nv, 128, 128, 8, 2, 2, 0])
    # rows a = 0, 1 (alpha) -> i = 0, 1; a = 2, 3 (beta) -> i = 8, 9
    rows = np.array([0, 1, 8, 9])
    np.testing.assert_array_equal(_h5dump_data(path, g + "hij_indx"), (rows[:, None] * 8 + np.arange(8)).ravel())
    nblk = (nnz + 99) // 100
    np.testing.assert_array_equal(_h5dump_data(path, "/Propagators/phaseless_ImpSamp_ForceBias/Spvn_dims"),
                                  [nnz, 64, 6, 8, nblk])


@pytest.mark.gpu
def test_driver_reads_hdf5_input(tmp_path, require_gpu):
    """The C++ driver given -f afqmc.h5 (the 64-orbital preset system written
    in the reference's format) prints the same step energies as the driver
    run on the preset generated in memory."""
    import re
    import subprocess
    p = m.PRESETS["N16_M64"]
    s = m.make_system(p["nmo"], p["naea"], p["nchol"], density=p["density"], sigma_L=p["sigma_L"], sigma_h=1.0,
                      dt=p["dt"], seed=p["seed"], with_vakbl=True)
    path = tmp_path / "afqmc.h5"
    hdf5.write_hdf5(s, path)
    exe = os.path.join(os.path.dirname(os.path.abspath(m.__file__)), "csrc", "miniafqmc")
    common = ["-i", "3", "-s", "2", "-o", "2", "-w", "16", "--dt", str(p["dt"]), "--seed", str(p["seed"])]

    def energies(args):
        out = subprocess.run([exe] + common + args, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        return np.array([float(v) for v in re.findall(r"^\d+\s+(-?\d+\.\d+)$", out.stdout, re.M)])

    e_file = energies(["-f", str(path)])
    e_mem = energies(["--system", "N16_M64"])
    assert len(e_file) == 3
    np.testing.assert_allclose(e_file, e_mem, rtol=1e-11)
