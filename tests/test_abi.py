"""The C-ABI library loads without a GPU and exports every declared symbol."""
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="afqmc_hip.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(afqmc_\w+)\s*\(", text)))


def test_library_exports_header():
    import miniafqmc_amd._native as nat
    syms = declared_symbols()
    assert len(syms) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", nat.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (afqmc_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    assert sorted(nat.EXPORTED) == syms
    for s in syms:
        getattr(nat.lib, s)


def test_h5_library_exports_header():
    """libafqmc_h5.so (built where the HDF5 C library is installed) exports
    every symbol afqmc_h5.h declares."""
    import pytest
    from miniafqmc_amd import hdf5
    if not hdf5.available():
        pytest.skip("libafqmc_h5.so not built (no HDF5 C library)")
    syms = declared_symbols("afqmc_h5.h")
    assert syms == ["afqmc_h5_destroy", "afqmc_h5_error", "afqmc_h5_get", "afqmc_h5_read", "afqmc_h5_write"]
    out = subprocess.run(["nm", "-D", "--defined-only", hdf5._PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (afqmc_\w+)", out))
    assert not [s for s in syms if s not in exported]


def test_ctx_create_rejects_bad_shapes_without_gpu():
    import ctypes as C
    import miniafqmc_amd._native as nat
    h = C.c_void_p()
    assert nat.lib.afqmc_ctx_create(0, 4, 8, 2, 0, C.byref(h)) == 2   # AFQMC_ERR_SHAPE
    assert nat.lib.afqmc_ctx_create(8, 9, 8, 2, 0, C.byref(h)) == 2   # N > M
    assert nat.lib.afqmc_ctx_create(8, 4, 8, 2, 7, C.byref(h)) == 1   # bad mode
    assert nat.lib.afqmc_ctx_create(600, 300, 8, 2, 0, C.byref(h)) == 8  # N > 256 unsupported
    assert nat.lib.afqmc_ctx_create(8, 4, 8, 2, 0, None) == 1


def test_synth_rejects_bad_params():
    import ctypes as C
    import miniafqmc_amd._native as nat
    h = C.c_void_p()
    p = nat.SynthParams(8, 10, 4, 0.5, 0.1, 1.0, 0.01, 1, 0)
    assert nat.lib.afqmc_synth_create(C.byref(p), C.byref(h)) == 2
    p = nat.SynthParams(8, 4, 4, 0.0, 0.1, 1.0, 0.01, 1, 0)
    assert nat.lib.afqmc_synth_create(C.byref(p), C.byref(h)) == 2


def test_product_imports_no_oracle():
    """The product package never references the oracle."""
    pkg = os.path.join(ROOT, "miniafqmc_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")) or f == "Makefile":
                txt = open(os.path.join(dp, f), errors="ignore").read()
                assert "oracle" not in txt.lower(), f


def test_python_constants_match_header():
    """Every status code, mode, buffer id and option of include/afqmc_hip.h
    has the same value in the Python mirror, and the mirror has no extra ones."""
    import miniafqmc_amd._native as nat
    text = open(os.path.join(ROOT, "include", "afqmc_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    enum = {k: int(v) for k, v in re.findall(r"\b(AFQMC_[A-Z0-9_]+)\s*=\s*(-?\d+)", text)}
    status = {k: v for k, v in enum.items() if k == "AFQMC_OK" or k.startswith("AFQMC_ERR_")}
    assert {v: k for k, v in status.items()} == nat.STATUS
    modes = {k[len("AFQMC_"):].lower(): v for k, v in enum.items()
             if k in ("AFQMC_FP64_CSR", "AFQMC_FP64_DENSE", "AFQMC_MIXED_FP32", "AFQMC_MIXED_BF16")}
    assert modes == nat.MODES
    for prefix in ("AFQMC_BUF_", "AFQMC_OPT_"):
        hdr = {k[len("AFQMC_"):]: v for k, v in enum.items() if k.startswith(prefix)}
        py = {k: getattr(nat, k) for k in dir(nat) if k.startswith(prefix[len("AFQMC_"):])}
        assert hdr == py, (prefix, hdr, py)


def test_noalloc_scratch_size_matches_header():
    """The Python mirror's size of the allocation-free entry point's scratch
    array is the header's AFQMC_NOALLOC_SCRATCH_BYTES."""
    import miniafqmc_amd._native as nat
    text = open(os.path.join(ROOT, "include", "afqmc_hip.h")).read()
    mm = re.search(r"#define\s+AFQMC_NOALLOC_SCRATCH_BYTES\s+\((\d+)u\s*<<\s*(\d+)\)", text)
    assert mm, "AFQMC_NOALLOC_SCRATCH_BYTES not found"
    assert int(mm.group(1)) << int(mm.group(2)) == nat.NOALLOC_SCRATCH_BYTES
