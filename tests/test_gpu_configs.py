"""The benchmark definition's larger configurations through the C-ABI against
the oracle, on a walker sample.

Walkers interact only through Eshift (the mean hybrid energy that scales the
weights, miniafqmc_base.cpp:317-327); their matrices, overlaps and hybrid
energies depend on their own fields alone, and the counter-based field
generator gives a walker the same draws in a run of any size.  So the GPU
runs every walker of the configuration and the oracle runs the first few;
those walkers must agree.

- 256 orbitals, 64+64 electrons, 1024 vectors at 20 % density, 256 walkers,
  FP64 CSR (the reference's algorithm), 5 steps of 10 substeps: walkers,
  overlaps and hybrid energies at 1e-9 relative after 50 substeps, and the
  local energy of walker 0 (Cholesky form) at 1e-9.
- 512 orbitals, 128+128 electrons, 2048 dense vectors, 1024 walkers: FP64
  dense at 1e-9 and split-BF16 at 1e-4 relative on walker 0 after two
  substeps (the production shapes: tiles, split bias GEMM, triangular
  exchange block, batched inverse at N = 128).
"""
import sys

import numpy as np
import pytest

import miniafqmc_amd as m
import oracle_lib as orc
from miniafqmc_amd import _native as nat

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def test_config_256_orbitals_fp64_csr(require_gpu):
    s, p = m.preset_system("N64_M256")
    nw, ns, nsub, northo = p["nwalk"], p["nsteps"], p["nsubsteps"], p["northo"]
    x = m.fields_block(p["seed"], s.nchol, nw, 0, ns * nsub)
    say(f"config 256 orbitals: GPU fp64_csr, {nw} walkers, {ns}x{nsub} substeps")
    g = m.AfqmcSys(s, nw, "fp64_csr")
    g.init_walkers()
    g.stage_fields(x)
    g.calculate_mixed_density_matrix()
    g.calculate_energy()
    g.run(ns, nsub, northo, s.dt)
    W, Wd = g.W, g.W_data
    g.close()
    ns_o = 3
    say(f"config 256 orbitals: oracle on walkers 0..{ns_o - 1}")
    spT = orc.halfrotate(s)
    xo = np.ascontiguousarray(x[:, :, :ns_o])
    eo, Wo, Wdo, eso, tm, info = orc.run(s, ns_o, ns, nsub, northo, xo, energy_mode=2, spT=spT)
    assert info == 0
    e = rel(W[:ns_o], Wo)
    say(f"walkers after {ns * nsub} substeps: rel err {e:.2e}")
    assert e < 1e-9
    assert rel(Wd[:ns_o, 2:4], Wdo[:, 2:4]) < 1e-9   # overlaps
    assert rel(Wd[:ns_o, 4:6], Wdo[:, 4:6]) < 1e-9   # hybrid energies
    # local energy of walker 0 on its final state (Cholesky form in the oracle)
    Wd0 = Wdo[:1].copy()
    G0, _ = orc.calculate_mixed_density_matrix(s, Wo[:1], Wd0)
    orc.energy_cholesky(s, Wd0, G0, orc.dense_Lt(s))
    say(f"E_L walker 0: GPU {Wd[0, 0].real:.10f} oracle {Wd0[0, 0].real:.10f}")
    assert abs(Wd[0, 0] - Wd0[0, 0]) <= 1e-9 * abs(Wd0[0, 0])


def test_config_512_orbitals_dense(require_gpu):
    s, p = m.preset_system("N128_M512")
    nw, nsub = p["nwalk"], 2
    x = m.fields_block(p["seed"], s.nchol, nw, 0, nsub)
    out = {}
    for mode in ("fp64_dense", "mixed_bf16"):
        say(f"config 512 orbitals: GPU {mode}, {nw} walkers, {nsub} substeps")
        g = m.AfqmcSys(s, nw, mode, dense_exchange=False)
        g.init_walkers()
        g.stage_fields(x)
        g.calculate_mixed_density_matrix()
        g.run(1, nsub, 10, s.dt, energy=False)
        out[mode] = g.get(nat.BUF_W)[:1]
        g.close()
    say("config 512 orbitals: oracle on walker 0")
    spT = orc.halfrotate(s, cutoff=-1.0)  # the dense modes keep every half-rotated entry
    eo, Wo, Wdo, eso, tm, info = orc.run(s, 1, 1, nsub, 10, np.ascontiguousarray(x[:, :, :1]), energy_mode=2,
                                         spT=spT)
    assert info == 0
    e64, e16 = rel(out["fp64_dense"], Wo), rel(out["mixed_bf16"], Wo)
    say(f"walker 0 after {nsub} substeps: fp64_dense rel err {e64:.2e}, mixed_bf16 {e16:.2e}")
    assert e64 < 1e-9
    assert e16 < 1e-4


# Stated bounds of the split-BF16 path at 512 orbitals / 128+128 electrons /
# 2048 vectors (energy level 2, exchange chunks of 2048, dense exchange
# block), one substep from the trial with shared fields, against the FP64
# oracle on walkers 0-1.  Relative errors are max |GPU - oracle| / max |oracle|
# over the walker's array.
MIXED_BOUNDS = {"G": 2e-5, "vbias": 2e-5, "X": 2e-5, "vHS": 2e-5, "hybridW": 2e-5, "W": 1e-4}
MIXED_DE_L = 1e-5  # Ha, |dE_L| target of the mixed path


@pytest.mark.timeout(900)
def test_config_512_orbitals_mixed_vs_oracle(require_gpu):
    """The benchmark's configuration (split-BF16 contractions and propagator,
    P W and P U with three-piece operands and the two innermost Horner steps
    with one-piece operands, FP64 LU/overlaps/energy sums,
    FP32-precision exchange in chunks of 2048, fields drawn on the device)
    and the same with two pieces throughout against the FP64 oracle, and the
    FP64 dense path against the oracle at 1e-10: one substep from the trial
    with the same draws (walkers 0-1 of the 1024), every phase of the
    substep and the local energy of the propagated walkers
    (afqmc_sys.hpp:95-133, vbias.hpp:45-83, vHS.hpp:43-91,
    energy.hpp:63-102)."""
    s, p = m.preset_system("N128_M512")
    nw, seed, K = p["nwalk"], p["seed"], s.nchol
    M, N = s.nmo, s.naea
    gpu = {}
    for key, mode, three in (("mixed_bf16", "mixed_bf16", 0), ("mixed_bf16 P W, P U three-piece", "mixed_bf16", 4),
                             ("fp64_dense", "fp64_dense", 0)):
        say(f"config 512 orbitals: GPU {key}, {nw} walkers, one substep from the trial")
        g = m.AfqmcSys(s, nw, mode, energy_precision=2, three_piece=three)
        if three:
            # the benchmark's default: P W and P U in three pieces, U_6 and
            # U_5 in one, the other Horner steps in two
            g.set_three_piece_steps(0x81)
            g.set_one_piece_steps(0x06)
        g.set_exchange_chunk(2048)
        g.device_fields(seed)
        g.init_walkers()
        g.calculate_mixed_density_matrix()
        g.substep(0, False, s.dt)
        r = {"G": g.get(nat.BUF_G)[:, :2], "vbias": g.get(nat.BUF_VBIAS)[:, :2], "X": g.get(nat.BUF_X)[:, :2],
             "vHS": g.get(nat.BUF_VHS)[:, :2], "hybridW": g.get(nat.BUF_HYBRIDW)[:2], "W": g.get(nat.BUF_W)[:2]}
        g.calculate_mixed_density_matrix()
        g.calculate_energy()
        r["E"] = g.W_data[:2, 0]
        r["terms"] = g.energy_terms()[:2]
        gpu[key] = r
        g.close()
    say("config 512 orbitals: oracle on walkers 0-1")
    W0 = np.zeros((2, 2, M, N), complex)
    W0[:, 0] = s.trial_alpha.conj()
    W0[:, 1] = s.trial_beta.conj()
    Wd = np.zeros((2, 8), complex)
    Wd[:, 1] = 1.0
    G0, info = orc.calculate_mixed_density_matrix(s, W0, Wd)
    assert info == 0
    spT = orc.halfrotate(s, cutoff=-1.0)  # dense modes keep every half-rotated entry
    vb = orc.get_vbias(spT, G0)
    x = m.fields(seed, K, 2, 0, w0=0, wtot=nw)
    X, hyb = orc.fields(vb, x)
    vhs = orc.get_vhs(s.spvn, X)
    W1 = orc.propagate(s, W0, vhs)
    Wd1 = Wd.copy()
    G1, info = orc.calculate_mixed_density_matrix(s, W1, Wd1)
    assert info == 0
    Lt = np.ascontiguousarray(spT.dense() / np.sqrt(s.dt))
    del spT
    say("config 512 orbitals: oracle local energy (Cholesky form)")
    orc.energy_cholesky(s, Wd1, G1, Lt)
    ref = {"G": G0, "vbias": vb, "X": X, "vHS": vhs, "hybridW": hyb, "W": W1}
    for mode, bounds in (("fp64_dense", None), ("mixed_bf16", MIXED_BOUNDS),
                         ("mixed_bf16 P W, P U three-piece", MIXED_BOUNDS)):
        r = gpu[mode]
        for k_, v in ref.items():
            for w in range(2):
                e = rel(r[k_][w] if k_ in ("hybridW", "W") else r[k_][:, w], v[w] if k_ in ("hybridW", "W") else v[:, w])
                say(f"{mode} walker {w} {k_}: rel err {e:.2e}")
                assert e < (1e-10 if bounds is None else bounds[k_]), (mode, k_, w, e)
        dE = np.abs(r["E"] - Wd1[:, 0])
        say(f"{mode} E_L: GPU {r['E'].real} oracle {Wd1[:, 0].real} |dE_L| {dE} Ha; terms {r['terms'].real}")
        if bounds is None:
            assert np.all(dE <= 1e-10 * np.abs(Wd1[:, 0]))
        else:
            assert np.all(dE < MIXED_DE_L)
