"""The C++ driver (miniafqmc_amd/csrc/miniafqmc, src/miniapps/miniafqmc_base.cpp
restated): with -v it reports the GPU time of the step loop under the
reference's timer names (miniafqmc_base.cpp:50-77)."""
import os
import re
import subprocess

import pytest

import miniafqmc_amd as m

pytestmark = pytest.mark.gpu


def test_driver_timer_report(require_gpu):
    exe = os.path.join(os.path.dirname(os.path.abspath(m.__file__)), "csrc", "miniafqmc")
    p = m.PRESETS["N16_M64"]
    args = ["-i", "2", "-s", "3", "-o", "2", "-w", "16", "-v", "--system", "N16_M64", "--dt", str(p["dt"]),
            "--seed", str(p["seed"])]
    print("driver -v N16_M64", flush=True)
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = {}
    for name, _ in m.REFERENCE_TIMERS:
        hit = re.search(r"^" + re.escape(name) + r"\s+([0-9.]+)\s+(\d+)$", out.stdout, re.M)
        assert hit, f"timer {name!r} missing from the report"
        rows[name] = (float(hit.group(1)), int(hit.group(2)))
    # 2 steps x 3 substeps: one vHS, one propagation (8 GEMM launches), one
    # set of fields per substep; one energy per step (plus the initial one)
    assert rows["H-S Potential"][1] >= 6 and rows["Propagation"][1] >= 6 * 8 and rows["Sigma"][1] >= 6
    assert rows["Local Energy"][1] > 0 and rows["Overlap"][1] >= 6
    assert all(t > 0 for t, n in rows.values() if n > 0)
    assert re.search(r"^Total \(wall\)\s+[0-9.]+", out.stdout, re.M)


def test_driver_branch_mixed_estimator(require_gpu):
    """--branch: comb population control after each step's energy
    (miniafqmc_base.cpp:346); the run ends normally and its mixed estimator
    is the mean and standard error of the printed step energies."""
    import numpy as np
    exe = os.path.join(os.path.dirname(os.path.abspath(m.__file__)), "csrc", "miniafqmc")
    p = m.PRESETS["N16_M64"]
    args = ["-i", "5", "-s", "4", "-o", "2", "-w", "32", "--branch", "--system", "N16_M64", "--dt", str(p["dt"]),
            "--seed", str(p["seed"])]
    print("driver --branch N16_M64", flush=True)
    out = subprocess.run([exe] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    e = np.array([float(v) for v in re.findall(r"^\d+\s+(-?\d+\.\d+)$", out.stdout, re.M)])
    assert len(e) == 5 and np.all(np.isfinite(e))
    hit = re.search(r"Mixed estimator over 5 steps: (-?[0-9.]+) \+/- ([0-9.eE+-]+)", out.stdout)
    assert hit, out.stdout
    mean, err = float(hit.group(1)), float(hit.group(2))
    assert abs(mean - e.mean()) < 1e-8 * abs(e.mean())
    # printed with four significant digits
    assert abs(err - e.std(ddof=1) / np.sqrt(len(e))) <= 1e-3 * err + 1e-9
