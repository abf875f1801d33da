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
    names |= set(re.findall(r'"([a-z_0-9]+_reduce)"', src))  # chunk sums named by their caller
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


def test_roofline_bound_follows_intensity():
    """The roofline entry takes the roof a call site's algorithmic work
    reaches last.  A Horner step at 512 orbitals / 128+128 electrons / 1024
    walkers streams V (2 MiB per walker in the split form) and U, S and U'
    (1 MiB each), plus U1's third pieces (0.5 MiB) once in six launches,
    for 8 * 2 * 512^2 * 128 flops per walker: about 101 flop/B, below the
    BF16 ridge (312), so HBM bounds it; the HS potential (681 flop/B) and
    the exchange stay MFMA-bound."""
    import argparse
    sys.path.insert(0, ROOT)
    import bench
    M, N, K, nw = 512, 128, 2048, 1024
    args = argparse.Namespace(mode="mixed_bf16", three_piece=4, three_steps=0x81, energy_level=2,
                              workload="N128_M512")
    stats = {"prop_taylor": (120, 120 * 1.6726), "vhs_gemm": (20, 20 * 9.56), "energy_vx_gemm": (64, 149.7),
             "fields": (20, 0.3)}
    roof, rows = bench.roofline_rows(stats, args, M, N, K, nw, n_energy=1)
    assert [r["kernel"] for r in rows] == ["prop_taylor", "vhs_gemm", "energy_vx_gemm"]
    mib = 1 << 20
    nbytes = nw * (2 + 3 + 0.5 / 6) * mib
    assert roof["kernel"] == "prop_taylor" and roof["bound"] == "hbm" and roof["unit"] == "GB/s"
    assert abs(roof["bytes_per_launch"] - nbytes) < 1.0
    assert abs(roof["achieved"] - nbytes / 1.6726e-3 / 1e9) < 0.1
    assert abs(roof["frac"] - nbytes / 1.6726e-3 / 8e12) < 1e-4
    flops = 8.0 * 2 * nw * M * M * N
    assert abs(roof["mfma_frac"] - flops / 1.6726e-3 / 2.5e15) < 1e-4
    by = {r["kernel"]: r for r in rows}
    assert by["vhs_gemm"]["bound"] == "mfma" and by["energy_vx_gemm"]["bound"] == "mfma"
    assert "hbm_frac" not in by["energy_vx_gemm"]
    # the FP32 mode reports the MFMA view only
    args.mode = "mixed_fp32"
    roof32, _ = bench.roofline_rows(stats, args, M, N, K, nw, n_energy=1)
    assert roof32["bound"] == "mfma"


def test_roofline_csr_sparse_product():
    """FP64 CSR mode: the sparse product with the most time is reported
    against HBM with its algorithmic bytes (16 B value + 4 B column per
    entry, the row pointers, the dense operand read once and the output
    written once) and its FP64 rate (one complex multiply-add per entry
    and walker)."""
    import argparse
    sys.path.insert(0, ROOT)
    import bench
    M, N, K, nw = 256, 64, 1024, 256
    nnz, nnzT = 13_400_000, 33_500_000
    args = argparse.Namespace(mode="fp64_csr", workload="N64_M256")
    csr = [("csrmm_vhs", (20, 20 * 2.355)), ("csrmm_vbias", (20, 20 * 1.0))]
    roof = bench.roofline_csr(csr, args, M, N, K, nw, nnz, nnzT)
    nbytes = nnz * 20 + (M * M + 1) * 4 + K * nw * 16 + M * M * nw * 16
    assert roof["kernel"] == "csrmm_vhs" and roof["bound"] == "hbm" and roof["bytes_per_launch"] == nbytes
    assert abs(roof["achieved"] - nbytes / 2.355e-3 / 1e9) < 0.1
    assert abs(roof["fp64_tflops"] - 8.0 * nnz * nw / 2.355e-3 / 1e12) < 0.01
