"""Walker-step throughput on one MI355X.

Workload (default): 512 orbitals, 128+128 electrons, 2048 dense Cholesky
vectors, 1024 walkers, dt = 0.005, synthetic Hamiltonian and auxiliary
fields (the reference ships no input file).  One step = one substep of all
walkers through mixed density matrix, vbias, fields, vHS, propagation,
overlaps and weights; the orthogonalisation runs every `--northo` substeps
and the local energy once at the end of the timed block, as the
reference's driver loop places them (miniafqmc_base.cpp:248-350).

Prints one JSON line: walker-steps/s of the chosen precision mode (default
split-BF16 contractions), the FP64 GPU path on the same workload and the
per-walker |dE_L| between the two, a roofline entry for the dominant kernel
(HIP events on the context's stream during the timed block), and the
oracle's CPU walker-steps/s on a walker sample.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {  # MI355X published dense peaks, TFLOP/s
    "fp64": 78.6,
    "fp32": 157.3,
    "bf16": 2500.0,
}
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
# the headline metric: the mixed-precision step at 512 orbitals, 128+128
# electrons and 2048 Cholesky vectors, compared with the FP64 path at a
# per-walker local-energy tolerance of 1e-5 Ha
HEADLINE_METRIC = "walker-steps/s at M=512,N=128,nchol=2048; mixed vs FP64 at |\u0394E_L|<1e-5 Ha"
MODE_PEAK = {"fp64_csr": "fp64", "fp64_dense": "fp64", "mixed_fp32": "fp32", "mixed_bf16": "bf16"}
MODE_DTYPE = {"fp64_csr": "c128", "fp64_dense": "c128", "mixed_fp32": "c64+c128", "mixed_bf16": "bf16x3+c128"}


# HBM bytes per launch measured with rocprofv3 (raw TCC read/write request
# counters, scaled by a 1 GiB copy), per (call site, workload, mode);
# profiles/walker_step_N128_M512_hbm_traffic.md has the commands and figures.
MEASURED_TRAFFIC = {
    # Horner steps, average of the six: 4.02 GiB read (V, U, S) + 1.09 GiB
    # written (U', and U1's third pieces) per launch
    ("prop_taylor", "N128_M512", "mixed_bf16"): 5.480e9,
}


def mixed_alg_bytes(name, M, N, K, nw, third_u1):
    """Algorithmic HBM bytes per launch of the split-BF16 call sites whose
    operands stream through once (8 B per complex in the split form, the
    same as a complex FP32): the Horner step reads V (M x M), U (M x 2N) and
    S as its FP32 addend, and writes U' (M x 2N); when P U takes three pieces
    the last of the six also writes U's third pieces (4 B per complex), so
    the average launch carries a sixth of that.  The HS potential reads the
    K Cholesky matrices and the fields and writes vHS."""
    if name == "prop_taylor":
        u = 8 * nw * M * 2 * N
        return 8 * nw * M * M + 3 * u + (u / 2 / 6 if third_u1 else 0)
    if name == "vhs_gemm":
        return 8 * K * M * M + 8 * K * nw + 8 * nw * M * M
    return None


def kernel_bytes(name, M, N, K, nw, nnz_spvn, nnz_spvnT):
    """Algorithmic HBM bytes per launch of the FP64 CSR products: the sparse
    matrix once (16 B value + 4 B column per entry, 4 B row pointers), the
    dense operand once and the output once (16 B per complex)."""
    if name == "csrmm_vhs":
        return nnz_spvn * 20 + (M * M + 1) * 4 + K * nw * 16 + M * M * nw * 16
    if name == "csrmm_vbias":
        return nnz_spvnT * 20 + (K + 1) * 4 + 2 * N * M * nw * 16 + K * nw * 16
    return None


def csr_flops(name, nw, nnz_spvn, nnz_spvnT):
    c = {"csrmm_vhs": nnz_spvn * nw, "csrmm_vbias": nnz_spvnT * nw}.get(name)
    return None if c is None else 8.0 * c


def contraction_spins_per_launch(M, N, K, nw):
    """Spin blocks per launch of the bias-shaped contraction (G against the
    K Cholesky rows, the bias and the Coulomb term).  The library splits its
    k = N*M range into S chunks, S doubling from 1 while 2*S*tiles < 2048
    (tiles = ceil(nw/128) * ceil(K/64)) and N*M/(2S) stays a multiple of 64;
    with S > 1 one launch covers both spins, otherwise each spin is its own
    launch."""
    NM = N * M
    tiles = -(-nw // 128) * -(-K // 64)
    S = 1
    while S < 16 and 2 * S * tiles < 2048 and NM % (2 * S) == 0 and (NM // (2 * S)) % 64 == 0:
        S *= 2
    return 2 if S > 1 else 1


def kernel_flops(name, M, N, K, nw, mode=None):
    """Algorithmic real flops per launch (8 per complex multiply-add).  The
    FP64 CSR mode's dense bias product contracts both spins in one launch."""
    sp = 2 if mode == "fp64_csr" else contraction_spins_per_launch(M, N, K, nw)
    c = {
        "vhs_gemm": nw * M * M * K,
        "prop_taylor": 2 * nw * M * M * N,
        "prop_pw": 2 * nw * M * M * N,
        "prop_ps": 2 * nw * M * M * N,
        "vbias_gemm": sp * nw * K * N * M,
        "energy_v_gemm": sp * nw * K * N * M,
        "ovlp_gemm": 2 * nw * N * N * M,
        "gc_gemm": 2 * nw * N * N * M,
        "gj_inverse": 2 * nw * N * N * N,
    }.get(name)
    return None if c is None else 8.0 * c


def kernel_flops_total(name, launches, M, N, K, nw, n_energy, mode=None):
    """Algorithmic real flops of all `launches` of a call site.  The exchange
    GEMM's launches differ in size (column chunks of the lower triangle), so
    its total is counted per energy evaluation: both spins, every walker,
    the NM(NM+1)/2 entries of the triangular block of the same-spin exchange
    matrix, one complex multiply-add each."""
    if name == "energy_vx_gemm":
        NM = N * M
        return 8.0 * n_energy * 2 * nw * NM * (NM + 1) / 2
    f = kernel_flops(name, M, N, K, nw, mode)
    return None if f is None else f * launches


def roofline_csr(csr, args, M, N, K, nw, nnz_spvn, nnz_spvnT):
    """Roofline entry of the FP64 CSR mode's sparse product with the most
    time (`csr`: (call site, (launches, ms)) pairs): HBM bandwidth against
    its algorithmic bytes, beside the FP64 rate it reaches."""
    name, (nl, ms) = max(csr, key=lambda kv: kv[1][1])
    avg_s = ms / nl / 1e3
    nbytes = kernel_bytes(name, M, N, K, nw, nnz_spvn, nnz_spvnT)
    fl = csr_flops(name, nw, nnz_spvn, nnz_spvnT)
    ach = nbytes / avg_s / 1e9
    return {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": MEASURED_TRAFFIC.get((name, args.workload, args.mode)),
            "avg_ms": round(ms / nl, 4), "bytes_per_launch": nbytes, "fp64_tflops": round(fl / avg_s / 1e12, 2)}


def roofline_rows(stats, args, M, N, K, nw, n_energy):
    """Roofline rows of the call sites with a known algorithmic count, most
    total time first, and the bench line's `roofline` entry for the first.
    A call site's bound is the roof its algorithmic work reaches last: HBM
    when its bytes at 8 TB/s take longer than its flops at the unit's peak
    (an intensity below the ridge point), the MFMA unit otherwise; both
    fractions are kept."""
    third_u1 = args.mode == "mixed_bf16" and bool(args.three_piece & 4) and bool(args.three_steps & 0x80)
    rows = []
    for k_, (nl, ms) in stats.items():
        fl = kernel_flops_total(k_, nl, M, N, K, nw, n_energy, args.mode)
        if fl is None or nl == 0 or ms <= 0:
            continue
        unit, pk = kernel_peak(k_, args.mode, args.energy_level, bool(getattr(args, "coulomb3", 0)))
        avg_s = ms / nl / 1e3
        ach = fl / (ms / 1e3) / 1e12
        r = {"kernel": k_, "achieved": round(ach, 2), "peak": pk, "peak_unit": unit,
             "frac": round(ach / pk, 4), "ms_total": round(ms, 2), "launches": nl,
             "avg_ms": round(ms / nl, 4), "flops_per_launch": fl / nl, "bound": "mfma", "unit": "TFLOP/s",
             "mfma_frac": round(ach / pk, 4), "tflops": ach, "mfma_unit": unit}
        nb = mixed_alg_bytes(k_, M, N, K, nw, third_u1) if args.mode == "mixed_bf16" else None
        if nb is not None:
            gbs = nb / avg_s / 1e9
            r["bytes_per_launch"] = nb
            r["hbm_frac"] = round(gbs / HBM_PEAK_GBS, 4)
            if nb / (HBM_PEAK_GBS * 1e9) > (fl / nl) / (pk * 1e12):
                r.update(bound="hbm", unit="GB/s", achieved=round(gbs, 1), peak=HBM_PEAK_GBS,
                         peak_unit="hbm", frac=r["hbm_frac"])
        rows.append(r)
    rows.sort(key=lambda r: -r["ms_total"])
    if not rows:
        return None, rows
    top = rows[0]
    roof = {"kernel": top["kernel"], "bound": top["bound"], "achieved": top["achieved"], "peak": top["peak"],
            "unit": top["unit"], "frac": top["frac"],
            "traffic": MEASURED_TRAFFIC.get((top["kernel"], args.workload, args.mode)),
            "avg_ms": top["avg_ms"], "flops_per_launch": top["flops_per_launch"],
            "bytes_per_launch": top.get("bytes_per_launch"), "mfma_frac": top["mfma_frac"],
            "hbm_frac": top.get("hbm_frac"),
            "kernels": [{k_: r[k_] for k_ in ("kernel", "bound", "achieved", "peak", "peak_unit", "frac",
                                              "mfma_frac", "ms_total", "launches")} for r in rows[:3]]}
    return roof, rows


def kernel_peak(name, mode, energy_level, coulomb3=False):
    """Peak (TFLOP/s) of the unit a call site runs on in a precision mode
    (coulomb3: the Coulomb contraction in three-piece split-BF16)."""
    if mode in ("fp64_csr", "fp64_dense") or name in ("ovlp_gemm", "gj_inverse"):
        return "fp64", PEAKS["fp64"]
    if name == "energy_vx_gemm":
        # level 2: three-piece split-BF16 in the split-BF16 mode, FP32 otherwise
        u = {0: MODE_PEAK[mode], 1: MODE_PEAK[mode], 2: "bf16" if mode == "mixed_bf16" else "fp32"}.get(
            energy_level, "fp64")
        return u, PEAKS[u]
    if name == "energy_v_gemm" and energy_level >= 1 and not (coulomb3 and mode == "mixed_bf16" and energy_level == 2):
        return "fp64", PEAKS["fp64"]
    return MODE_PEAK[mode], PEAKS[MODE_PEAK[mode]]


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else avail
    return model, avail, threads


def bench_multi(args, m):
    """Uneven batch of independent systems (M = 96 ... 512) advanced together,
    one graph launch per substep, against the same systems run one after
    another.  value = walkers x substeps summed over the systems / time."""
    t0 = time.time()
    systems = m.multi_systems()
    print(f"[bench] {len(systems)} systems generated in {time.time() - t0:.1f}s", flush=True)
    nsub = args.warmup + args.steps
    spe = args.substeps_per_energy or args.steps
    northo = args.northo or 10

    def contexts():
        out = []
        for s, nw, seed in systems:
            g = m.AfqmcSys(s, nw, args.mode, energy_precision=args.energy_level)
            g.set_exchange_chunk(args.exchange_chunk)
            g.set_direct_lds(args.direct_lds)
            g.stage_fields(m.fields_block(seed, s.nchol, nw, 0, nsub))
            if args.tune:
                g.tune(s.dt)
            g.init_walkers()
            g.calculate_mixed_density_matrix()
            g.calculate_energy()
            out.append(g)
        return out

    def block(step_fn, ctxs):
        for t in range(args.warmup):
            step_fn(t)
        for g in ctxs:
            g.sync()
        ta = time.perf_counter()
        for t in range(args.warmup, nsub):
            step_fn(t)
            if (t - args.warmup + 1) % spe == 0:
                for g in ctxs:
                    g.calculate_mixed_density_matrix()
                    g.calculate_energy()
        for g in ctxs:
            g.sync()
        return time.perf_counter() - ta

    total_w = sum(nw for _, nw, _ in systems)
    ctxs = contexts()
    grp = m.AfqmcGroup(ctxs)
    Tg = block(lambda t: grp.substep(t, t > 0 and t % northo == 0, 0.005), ctxs)
    grp.close()
    for g in ctxs:
        g.close()
    ctxs = contexts()
    if args.graphs:
        for g in ctxs:
            g.use_graphs(True)

    def seq(t):
        for g in ctxs:
            g.substep(t, t > 0 and t % northo == 0, 0.005)
    Ts = block(seq, ctxs)
    for g in ctxs:
        g.close()
    value = total_w * args.steps / Tg
    vseq = total_w * args.steps / Ts
    print(f"[bench] group: {Tg:.3f}s, one after another: {Ts:.3f}s for {args.steps} substeps", flush=True)
    line = {
        "metric": f"walker-steps/s ({len(systems)} uneven systems, M={systems[0][0].nmo}..{systems[-1][0].nmo}, "
                  f"{total_w} walkers in total)",
        "value": round(value, 2), "unit": "walker-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * Tg / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": MODE_DTYPE[args.mode],
        "data": "synthetic (seeded Hamiltonians and auxiliary fields)",
        "config": {"workload": "multi", "systems": [[s.nmo, s.naea, s.nchol, nw] for s, nw, _ in systems],
                   "mode": args.mode, "northo": northo, "energy_level": args.energy_level,
                   "energy": f"once per {spe} substeps", "parallelism": "single GPU, one graph launch per substep"},
        "one_after_another": {"value": round(vseq, 2), "graphs": bool(args.graphs)},
        "speedup_vs_one_after_another": round(value / vseq, 3),
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="N128_M512")
    ap.add_argument("--mode", default="mixed_bf16")
    ap.add_argument("--northo", type=int, default=None)
    ap.add_argument("--walkers", type=int, default=None)
    ap.add_argument("--fp64", type=int, default=1, help="also run the FP64 GPU path (0 to skip)")
    ap.add_argument("--cpu", type=int, default=1, help="time the oracle on a walker sample (0 to skip)")
    ap.add_argument("--cpu-walkers", type=int, default=6)
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--energy-level", type=int, default=2,
                    help="mixed modes: 0 energy in the contraction precision, 1 FP64 G/one-body/Coulomb, "
                         "2 and FP32 exchange (FP64 sums over column chunks), 3 and FP64 exchange")
    ap.add_argument("--exchange-chunk", type=int, default=2048,
                    help="mixed levels 0-2: exchange columns per FP32-accumulated chunk (0: one pass)")
    ap.add_argument("--substeps-per-energy", type=int, default=0,
                    help="substeps per driver step (the local energy follows each step); 0: one step "
                         "of all timed substeps")
    ap.add_argument("--tune", type=int, default=1, help="autotune GEMM tiles before the warmup (untimed)")
    ap.add_argument("--graphs", type=int, default=1, help="replay substeps from captured HIP graphs (0: eager)")
    ap.add_argument("--direct-lds", type=int, default=2,
                    help="split-BF16 GEMMs: 2 both split operands through a ring of LDS stages by direct loads, "
                         "1 the left operand by direct loads into two stages, 0 through registers")
    ap.add_argument("--three-piece", type=int, default=4,
                    help="split-BF16 mode: contractions in three-piece split-BF16 (24-bit operands), a mask: "
                         "1 bias product, 2 HS potential, 4 propagator")
    ap.add_argument("--three-steps", type=lambda v: int(v, 0), default=0x81,
                    help="with the propagator in three pieces: a mask of its eight products that take three "
                         "pieces (bit 0 P W, bits 1-6 the Horner steps U_6 .. U_1, bit 7 P U)")
    ap.add_argument("--one-steps", type=lambda v: int(v, 0), default=0x06,
                    help="with the propagator in three pieces: a mask of the Horner steps (bits 1-6: U_6 .. U_1) "
                         "whose products take one piece (the hi bf16 words)")
    ap.add_argument("--coulomb3", type=int, default=0,
                    help="split-BF16 mode at energy level 2: the local energy's Coulomb contraction in three-piece "
                         "split-BF16 instead of FP64")
    ap.add_argument("--variant", type=int, default=1,
                    help="split-BF16 mode: also time the other piece count (two pieces if --three-piece is set, "
                         "else all three-piece) from the same walkers and report it beside the value")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run one device-to-device copy of this many MiB and exit (reference for "
                         "hardware byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; the walker-sharded multi-GPU "
              "path is a design note, not built (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib > 0:
        # reference for hardware byte counters: one device-to-device copy of
        # known size through the HIP runtime alone, in a process of its own
        # (no other GPU library is loaded; the fill before it is a separate
        # dispatch)
        hip = C.CDLL("libamdhip64.so")
        nbytes = C.c_size_t(args.calib_copy_mib << 20)
        src, dst = C.c_void_p(), C.c_void_p()
        for ptr in (src, dst):
            if hip.hipMalloc(C.byref(ptr), nbytes) != 0:
                sys.exit("bench.py: hipMalloc failed")
        ok = (hip.hipMemset(src, 1, nbytes) == 0 and hip.hipDeviceSynchronize() == 0
              and hip.hipMemcpy(dst, src, nbytes, 3) == 0 and hip.hipDeviceSynchronize() == 0)  # 3: device to device
        hip.hipFree(src)
        hip.hipFree(dst)
        if not ok:
            sys.exit("bench.py: calibration copy failed")
        print(f"[bench] calibration copy: {args.calib_copy_mib} MiB read + {args.calib_copy_mib} MiB written",
              flush=True)
        return

    import miniafqmc_amd as m
    from miniafqmc_amd import _native as nat

    if args.workload == "multi":
        return bench_multi(args, m)


    t0 = time.time()
    s, p = m.preset_system(args.workload)
    nw = args.walkers or p["nwalk"]
    northo = args.northo or p["northo"]
    M, N, K = s.nmo, s.naea, s.nchol
    t_gen = time.time() - t0
    nsub = args.warmup + args.steps
    spe = args.substeps_per_energy or args.steps
    if args.steps % spe:
        sys.exit("bench.py: --steps must be a multiple of --substeps-per-energy")
    print(f"[bench] system M={M} N={N} K={K} nnz={s.spvn.nnz} walkers={nw} generated in {t_gen:.1f}s", flush=True)

    probe_steps = (1, 10)

    def probe(g, state):
        """Untimed accuracy probes from a given walker state (W, W_data, Eshift):
        the local energy of those walkers, then of the walkers after k substeps
        with the same fields, for k in probe_steps."""
        W0, Wd0, es0 = state
        out = {}
        g.set(nat.BUF_W, W0)
        g.set(nat.BUF_WDATA, Wd0)
        g.calculate_mixed_density_matrix()
        g.calculate_energy()
        out["eval"] = (g.W_data[:, 0], g.energy_terms())
        for k in probe_steps:
            g.set(nat.BUF_W, W0)
            g.set(nat.BUF_WDATA, Wd0)
            g.set_eshift(es0)
            for t in range(k):
                g.substep(t, t > 0 and t % northo == 0, s.dt)
            g.calculate_mixed_density_matrix()
            g.calculate_energy()
            out[k] = g.W_data[:, 0]
        return out

    def run_mode(mode, profile, state=None, three=None):
        """Time `args.steps` substeps in `mode`; then run the accuracy probes
        from `state`, or from this run's own final walkers (returned)."""
        three = args.three_piece if three is None else three
        g = m.AfqmcSys(s, nw, mode, energy_precision=args.energy_level,
                       three_piece=three if mode == "mixed_bf16" else 0,
                       coulomb_three_piece=bool(args.coulomb3) and mode == "mixed_bf16")
        g.set_exchange_chunk(args.exchange_chunk)
        g.set_direct_lds(args.direct_lds)
        if mode == "mixed_bf16" and three & 4:
            g.set_three_piece_steps(args.three_steps)
            g.set_one_piece_steps(args.one_steps)
        # the auxiliary fields are drawn on the device inside each substep
        # (the reference's generate_normal in the step loop), the same
        # counter-based draws the oracle is given on the host
        g.device_fields(p["seed"])
        if args.tune:
            tt = time.perf_counter()
            g.tune(s.dt)
            print(f"[bench] {mode}: tile autotune {time.perf_counter() - tt:.1f}s (untimed)", flush=True)
        g.init_walkers()
        if args.graphs:
            g.use_graphs(True)
        g.calculate_mixed_density_matrix()
        if not args.no_energy:
            g.calculate_energy()
        for t in range(args.warmup):
            g.substep(t, t > 0 and t % northo == 0, s.dt)
        g.sync()
        # HIP-event kernel timing runs eagerly: with graphs the timed block
        # replays graphs without events, and a profiled eager block of the
        # same substeps follows it (untimed for the value)
        eager_prof = profile and not args.graphs
        if eager_prof:
            g.profile(True)
        ta = time.perf_counter()
        eav = g.run(args.steps // spe, spe, northo, s.dt, first=args.warmup, energy=not args.no_energy)
        g.sync()
        tb = time.perf_counter()
        out = {"Wd": g.W_data}
        if mode == "fp64_csr":
            nnz = C.c_longlong(0)
            nat.check(nat.lib.afqmc_get_spvnT(g.ctx, C.byref(nnz), None, None, None))
            out["nnz_spvnT"] = nnz.value
        if profile and args.graphs:
            st = (g.W, out["Wd"], g.eshift())
            g.profile(True)
            g.run(args.steps // spe, spe, northo, s.dt, first=args.warmup, energy=not args.no_energy)
            g.sync()
            # restore the timed run's end state for the probes below
            g.set(nat.BUF_W, st[0])
            g.set(nat.BUF_WDATA, st[1])
            g.set_eshift(st[2])
        stats = g.kernel_stats() if profile else {}
        if profile:
            g.profile(False)
        out["stats"] = stats
        if not args.no_energy and args.fp64:
            if state is None:
                state = (g.W, out["Wd"], g.eshift())
                out["state"] = state
            out["probe"] = probe(g, state)
        g.close()
        return tb - ta, eav, out

    T, eav, res = run_mode(args.mode, True)
    Wd, stats = res["Wd"], res["stats"]
    value = nw * args.steps / T
    print(f"[bench] {args.mode}: {T:.3f}s for {args.steps} substeps -> {value:.1f} walker-steps/s", flush=True)
    for k_, (n_, ms) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        print(f"[bench]   {k_:20s} {n_:6d} launches {ms:10.2f} ms total {ms / max(n_, 1):9.3f} ms avg", flush=True)

    # roofline of the dominant kernel with a known algorithmic count: the
    # call site with the largest total time in the timed block among those
    # whose algorithmic work is known; the next two follow in "kernels"
    roof = None
    n_energy = 0 if args.no_energy else args.steps // spe
    csr = [(k_, v) for k_, v in stats.items() if k_ in ("csrmm_vhs", "csrmm_vbias")]
    if args.mode == "fp64_csr" and csr:
        roof = roofline_csr(csr, args, M, N, K, nw, s.spvn.nnz, res["nnz_spvnT"])
    else:
        roof, rows = roofline_rows(stats, args, M, N, K, nw, n_energy)
        for r in rows:
            hb = f", {r['hbm_frac']:.3f} of HBM" if "hbm_frac" in r else ""
            print(f"[bench] roofline {r['kernel']:16s} bound {r['bound']:4s} {r['tflops']:8.1f} TFLOP/s = "
                  f"{r['mfma_frac']:.3f} of {r['mfma_unit']}{hb}; {r['ms_total']:.1f} ms", flush=True)

    extra = {}
    if stats:
        # the same kernel time under the reference driver's timer names
        # (src/miniapps/miniafqmc_base.cpp:50-77); Total is the timed block
        rt = m.reference_timers(stats)
        extra["reference_timers_ms"] = {"Total": round(1e3 * T, 2), **{k_: round(ms, 2) for k_, (n_, ms) in rt.items()}}
        print("[bench] " + ", ".join(f"{k_} {v_:.1f} ms" for k_, v_ in extra["reference_timers_ms"].items()), flush=True)
    if args.fp64 and args.mode != "fp64_dense":
        T64, eav64, r64 = run_mode("fp64_dense", True, state=res.get("state"))
        Wd64 = r64["Wd"]
        v64 = nw * args.steps / T64
        extra["fp64_gpu"] = {"value": round(v64, 2), "ms_per_step": round(1e3 * T64 / args.steps, 3),
                             "kernels_ms": {k_: round(ms, 2) for k_, (n_, ms) in r64["stats"].items() if ms > 1.0}}
        extra["speedup_vs_fp64"] = round(value / v64, 3)
        if not args.no_energy:
            pm, p64 = res["probe"], r64["probe"]
            e64 = p64["eval"][0]
            Eref = float(np.abs(e64).mean())

            def err(a, b):
                d = np.abs(a - b)
                return {"max_abs_dE_L_Ha": float(d.max()), "mean_abs_dE_L_Ha": float(d.mean()),
                        "max_rel": float(d.max() / Eref)}
            # (1) energy evaluation: both precisions on the same walkers
            ev = err(pm["eval"][0], e64)
            ev["max_abs_dterm_Ha"] = dict(zip(("one_body", "coulomb", "exchange"),
                                              map(float, np.abs(pm["eval"][1] - p64["eval"][1]).max(axis=0))))
            ev["E_L_terms_mean_Ha"] = [float(v) for v in p64["eval"][1].real.mean(axis=0)]
            extra["energy_eval_error"] = ev
            # (2) k substeps from the same walkers and fields, each in its own precision
            extra["step_error"] = {f"{k}_substeps": err(pm[k], p64[k]) for k in probe_steps}
            extra["target_abs_dE_L_Ha"] = 1e-5
            # (3) the two timed runs' own walkers after the whole block
            dE = np.abs(Wd[:, 0] - Wd64[:, 0])
            extra["block_trajectory_error"] = {
                "max_abs_dE_L_Ha": float(np.nanmax(dE)), "mean_abs_dE_L_Ha": float(np.nanmean(dE)),
                "Eav_Ha": [float(eav[-1]), float(eav64[-1])], "substeps": args.warmup + args.steps,
            }
            extra["E_L_mean_Ha"] = float(Wd64[:, 0].real.mean())
            # the other split-BF16 piece count on the same workload, from the
            # same walkers: its speed and its errors against the same FP64 probes
            if args.mode == "mixed_bf16" and args.variant:
                alt = 0 if args.three_piece else 7
                Tv, _, rv = run_mode(args.mode, False, state=res.get("state"), three=alt)
                vv = nw * args.steps / Tv
                extra["variant"] = {
                    "three_piece": alt, "value": round(vv, 2), "ms_per_step": round(1e3 * Tv / args.steps, 3),
                    "speedup_vs_fp64": round(vv / v64, 3),
                    "energy_eval_error": err(rv["probe"]["eval"][0], e64),
                    "step_error": {f"{k}_substeps": err(rv["probe"][k], p64[k]) for k in probe_steps}}
                print(f"[bench] variant three_piece={alt}: {vv:.1f} walker-steps/s ({vv / v64:.2f}x fp64); "
                      f"step_error {extra['variant']['step_error']}", flush=True)
        print(f"[bench] fp64_dense: {v64:.1f} walker-steps/s; mixed/fp64 = {value / v64:.2f}x", flush=True)
        for k_, (n_, ms) in sorted(r64["stats"].items(), key=lambda kv: -kv[1][1])[:12]:
            print(f"[bench]   fp64 {k_:20s} {n_:6d} launches {ms:10.2f} ms total {ms / max(n_, 1):9.3f} ms avg",
                  flush=True)
        for key in ("energy_eval_error", "step_error", "block_trajectory_error"):
            print(f"[bench] {key}: {extra.get(key)}", flush=True)

    cpu = None
    if args.cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        model, avail, threads = cpu_info()
        tr = time.time()
        orc.set_threads(threads)
        spT = orc.halfrotate(s, cutoff=1e-6)
        print(f"[bench] oracle half rotation {time.time() - tr:.1f}s on {threads} threads (setup, untimed)",
              flush=True)

        def cpu_leg(cw, nthreads):
            # walkers 0..cw-1 of the same run: the counter-based draws give
            # them the fields they have in the full GPU run
            xs = m.fields(p["seed"], K, cw, 0, w0=0, wtot=nw)[None].copy()
            orc.set_threads(nthreads)
            eo, Wo, Wdo, eso, tm, info = orc.run(s, cw, 1, 1, 1 << 30, xs, energy_mode=2, spT=spT)
            return tm[0]

        cw1, cwa = args.cpu_walkers, max(args.cpu_walkers, 4 * threads)
        t1 = cpu_leg(cw1, 1)
        ta_ = cpu_leg(cwa, threads)
        orc.set_threads(threads)
        cpu = {"value": round(cwa / ta_, 4), "unit": "walker-steps/s", "cores": threads, "kind": "port",
               "cpu_model": model, "cpus_available": avail,
               "sample": f"{cwa} walkers x 1 substep of the same workload (no energy), {ta_:.1f}s on {threads} "
                         f"threads (OpenMP over walkers); oracle = FP64 C++ restatement of the reference, "
                         f"CSR SpvnT/Spvn",
               "single_thread": {"value": round(cw1 / t1, 4), "cores": 1,
                                 "sample": f"{cw1} walkers x 1 substep, {t1:.1f}s on 1 thread (the reference's "
                                           f"serial execution model)"}}
        print(f"[bench] cpu oracle: {cw1 / t1:.3f} walker-steps/s on 1 thread, {cwa / ta_:.3f} on {threads} "
              f"threads ({model})", flush=True)

    line = {
        "metric": (HEADLINE_METRIC if args.workload == "N128_M512" and args.mode == "mixed_bf16"
                   else f"walker-steps/s (M={M}, N={N}+{N}, nchol={K}, {nw} walkers, {args.mode})"),
        "value": round(value, 2),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * T / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # split-BF16 with some contractions in three pieces (six MFMAs a product)
        "dtype": "bf16x3+bf16x6+c128" if args.mode == "mixed_bf16" and args.three_piece else MODE_DTYPE[args.mode],
        "data": "synthetic (seeded Hamiltonian and auxiliary fields)",
        "config": {"workload": args.workload, "nmo": M, "naea": N, "nchol": K, "walkers": nw, "dt": s.dt,
                   "mode": args.mode, "northo": northo, "energy_level": args.energy_level,
                   "exchange_chunk": args.exchange_chunk if args.energy_level < 3 else None,
                   "graphs": bool(args.graphs), "direct_lds": args.direct_lds,
                   "three_piece": args.three_piece if args.mode == "mixed_bf16" else 0,
                   "three_piece_steps": args.three_steps if args.mode == "mixed_bf16" and args.three_piece & 4
                   else None,
                   "one_piece_steps": args.one_steps if args.mode == "mixed_bf16" and args.three_piece & 4
                   else None,
                   "coulomb": ("three-piece split-BF16" if args.coulomb3 and args.mode == "mixed_bf16"
                               and args.energy_level == 2 else "FP64" if args.energy_level >= 1 else "contraction"),
                   "energy": "none" if args.no_energy else f"once per {spe} substeps",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(extra)

    def finite(v):  # strict JSON: non-finite numbers become null
        if isinstance(v, dict):
            return {k_: finite(x_) for k_, x_ in v.items()}
        if isinstance(v, (list, tuple)):
            return [finite(x_) for x_ in v]
        if isinstance(v, float) and not math.isfinite(v):
            return None
        return v
    print(json.dumps(finite(line)), flush=True)


if __name__ == "__main__":
    main()
