"""bench.py runs on one GPU only: asked for more (by flag or by a launcher's
WORLD_SIZE) it exits with status 2 and says so, before touching a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True, text=True,
                          env=env, timeout=120)


def test_bench_refuses_several_gpus_by_flag():
    r = run_bench(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_bench_refuses_several_ranks_by_environment():
    r = run_bench(["--steps", "1", "--warmup", "0"], {"WORLD_SIZE": "2", "RANK": "0", "LOCAL_RANK": "0"})
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_reference_timers_cover_every_call_site():
    """Every call site the library launches under during a substep or an
    energy is summed under one of the reference driver's timers
    (src/miniapps/miniafqmc_base.cpp:50-77); the others run at upload or in
    the public GEMM entry point."""
    import re
    import miniafqmc_amd as m
    src = open(os.path.join(ROOT, "miniafqmc_amd", "csrc", "ctx.hip")).read()
    names = set(re.findall(r'launch\(c, "([a-z_0-9]+)"', src))
    names |= set(re.findall(r'\(c, "([a-z_0-9]+)", a', src))
    names |= set(re.findall(r'csrmm_n\(c, "([a-z_0-9]+)"', src))
    setup = {"build_vx", "vx_lower", "vx_split3", "halfrotate_gemm", "rot_count", "rot_fill", "csr_densify",
             "third_pieces", "init_walkers", "wdata_logs", "zgemm"}
    timed = [k for _, sites in m.REFERENCE_TIMERS for k in sites]
    assert len(timed) == len(set(timed)), "a call site under two timers"
    missing = sorted(names - set(timed) - setup)
    assert not missing, missing
    assert [t for t, _ in m.REFERENCE_TIMERS] == [
        "compact Mixed Density Matrix", "Mixed Density Matrix", "Bias Potential", "H-S Potential", "Sigma",
        "Propagation", "Other", "Overlap", "Orthgonalization", "Local Energy"]
    got = m.reference_timers({"prop_taylor": (6, 12.0), "prop_pw": (1, 2.0), "gj_inverse": (1, 5.0), "x": (1, 1.0)})
    assert got["Propagation"] == (7, 14.0) and got["Overlap"] == (1, 5.0) and got["Sigma"] == (0, 0)
