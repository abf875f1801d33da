"""HIP path vs the oracle through the C-ABI, phase by phase and end to end.

FP64 bar: 1e-10 relative (per element, relative to the array's max
magnitude) for every phase; the multi-step trajectory compares at 1e-9.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import miniafqmc_amd as m
import oracle_lib as orc
from miniafqmc_amd import _native as nat

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def close(a, b, rtol=RTOL, what=""):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f"{what}: max rel err {err:.3e} > {rtol:.1e}"
    return err


@pytest.fixture(scope="module")
def c1(require_gpu):
    s, p = m.preset_system("N16_M64")
    return s, p


def perturbed_walkers(s, nw, seed=3, noise=0.1):
    r = np.random.default_rng(seed)
    W = np.zeros((nw, 2, s.nmo, s.naea), complex)
    W[:, 0] = s.trial_alpha.conj()
    W[:, 1] = s.trial_beta.conj()
    W += noise * (r.standard_normal(W.shape) + 1j * r.standard_normal(W.shape))
    return W


@pytest.mark.parametrize("M,N,nw", [(512, 128, 64), (200, 64, 32), (40, 16, 16)])
def test_mixed_density_matrix_row_swaps(require_gpu, M, N, nw):
    """Walkers unrelated to the trial (complex Gaussian entries): partial
    pivoting swaps rows at most steps of the batched inverse, many of them
    rows held by different waves (the case where the swapped-in row must
    read old row k's column entry from this step's copy).  G and the
    overlaps against the oracle at 1e-10."""
    say("mdm with row swaps", M, N, nw)
    s = m.make_system(M, N, 8, density=1.0, sigma_L=0.01, sigma_h=1.0, dt=0.005, seed=17 + M)
    r = np.random.default_rng(M + N)
    W = r.standard_normal((nw, 2, M, N)) + 1j * r.standard_normal((nw, 2, M, N))
    Wd = np.zeros((nw, 8), complex)
    g = m.AfqmcSys(s, nw, "fp64_dense")
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, Wd)
    G = g.calculate_mixed_density_matrix()
    ov = g.W_data[:, 2:4].copy()
    g.close()
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
    # per walker: the overlaps of unrelated walkers span many decades
    eg = (np.abs(G - Go).max(axis=0) / np.abs(Go).max(axis=0)).max()
    eo = (np.abs(ov - Wd[:, 2:4]) / np.abs(Wd[:, 2:4])).max()
    say(f"  G per-walker rel err {eg:.2e}, overlaps {eo:.2e}")
    assert eg < 1e-10 and eo < 1e-10


def test_halfrotate(c1):
    s, p = c1
    say("halfrotate N16_M64")
    g = m.AfqmcSys(s, 4, "fp64_csr")
    spT_g = g.spvnT()
    spT_o = orc.halfrotate(s, cutoff=1e-6)
    Dg, Do = spT_g.dense(), spT_o.dense()
    both = (Dg != 0) & (Do != 0)
    close(Dg[both], Do[both], 1e-12, "SpvnT values")
    only = (Dg != 0) ^ (Do != 0)
    assert np.all(np.abs(np.abs(np.where(only, Dg + Do, 0))[only] - 1e-6) < 1e-13)
    assert abs(spT_g.nnz - spT_o.nnz) <= 2
    for r in range(s.nchol):
        c = spT_g.col[spT_g.rowptr[r]:spT_g.rowptr[r + 1]]
        assert np.all(np.diff(c) > 0)
    g.close()


@pytest.mark.parametrize("mode", ["fp64_csr", "fp64_dense"])
def test_mixed_density_matrix(c1, mode):
    s, p = c1
    nw = p["nwalk"]
    say("mdm", mode)
    g = m.AfqmcSys(s, nw, mode)
    W = perturbed_walkers(s, nw)
    Wd = np.zeros((nw, 8), complex)
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, Wd)
    G = g.calculate_mixed_density_matrix()
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
    close(G, Go, what="G")
    close(g.W_data[:, 2:4], Wd[:, 2:4], what="overlaps")
    g.close()


def test_phases_fp64_csr(c1):
    """vbias, fields, vHS, propagate, overlaps, weights, ortho, energy with shared inputs."""
    s, p = c1
    nw = p["nwalk"]
    g = m.AfqmcSys(s, nw, "fp64_csr")
    spT = orc.halfrotate(s)
    W = perturbed_walkers(s, nw, seed=5)
    Wd = np.zeros((nw, 8), complex)
    Wd[:, 1] = 1.0
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, Wd)
    say("phases: mdm")
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
    g.calculate_mixed_density_matrix()
    say("phases: vbias")
    g.set(nat.BUF_G, Go)
    vb = g.get_vbias()
    vbo = orc.get_vbias(spT, Go)
    close(vb, vbo, what="vbias")
    say("phases: fields")
    x = m.fields(p["seed"], s.nchol, nw, 0)
    g.set(nat.BUF_VBIAS, vbo)
    g.fields(x)
    Xo, ho = orc.fields(vbo, x)
    close(g.get(nat.BUF_X), Xo, what="X")
    close(g.get(nat.BUF_HYBRIDW), ho, what="hybridW")
    say("phases: vHS")
    g.set(nat.BUF_X, Xo)
    vhs = g.get_vHS()
    vhso = orc.get_vhs(s.spvn, Xo)
    close(vhs, vhso, what="vHS")
    say("phases: propagate")
    g.set(nat.BUF_VHS, vhso)
    g.propagate()
    Wo = orc.propagate(s, W, vhso)
    close(g.W, Wo, what="propagated W")
    say("phases: overlaps + weights")
    g.set(nat.BUF_W, Wo)
    g.save_overlaps()
    g.calculate_overlaps()
    Wd[:, 5] = Wd[:, 4]
    Wd[:, 6:8] = Wd[:, 2:4]
    orc.calculate_overlaps(s, Wo, Wd)
    close(g.W_data[:, 2:4], Wd[:, 2:4], what="overlaps")
    g.set(nat.BUF_HYBRIDW, ho)
    es = g.weights(s.dt, 0.25)
    eso = orc.weights(s.dt, ho, Wd, 0.25)
    close(g.W_data, Wd, what="W_data after weights")
    assert abs(es - eso) <= 1e-10 * abs(eso)
    say("phases: orthogonalize")
    g.orthogonalize()
    Woo = orc.orthogonalize(Wo)
    close(g.W, Woo, what="orthogonalized W")
    say("phases: energy")
    g.set(nat.BUF_WDATA, Wd)
    g.calculate_mixed_density_matrix()
    e = g.calculate_energy()
    G2, _ = orc.calculate_mixed_density_matrix(s, Woo, Wd)
    eo = orc.calculate_energy(s, Wd, G2)
    close(g.W_data[:, 0], Wd[:, 0], what="E_L (Vakbl)")
    assert abs(e - eo) <= 1e-10 * abs(eo)
    g.close()


@pytest.mark.parametrize("householder", [False, True])
@pytest.mark.parametrize("M,N", [(100, 24), (300, 100), (512, 128)])
def test_orthogonalize_sizes(require_gpu, M, N, householder):
    """LAPACK LQ factor (Cholesky QR + sign reconstruction, or the blocked
    Householder kernel with ragged panels) up to 512 x 128 vs the oracle's
    zgelq2/zungl2, on Gaussian matrices and on an ill-conditioned one that
    the Cholesky route hands to the Householder kernel."""
    say("orthogonalize", M, N, "householder" if householder else "cholesky qr")
    s = m.make_system(M, N, 2, density=1.0, sigma_L=0.01, seed=M)
    nw = 3
    g = m.AfqmcSys(s, nw, "fp64_csr", dense_exchange=False)
    g.set_ortho_householder(householder)
    r = np.random.default_rng(M)
    W = r.standard_normal((nw, 2, M, N)) + 1j * r.standard_normal((nw, 2, M, N))
    # nearly parallel columns (condition ~1e5): the Cholesky route declines
    # this matrix and the Householder kernel takes it
    W[1, 0, :, 1] = W[1, 0, :, 0] + 5e-5 * (r.standard_normal(M) + 1j * r.standard_normal(M))
    g.set(nat.BUF_W, W)
    g.orthogonalize()
    close(g.W, orc.orthogonalize(W), 1e-10, "orthogonalized walkers")
    g.close()


@pytest.mark.parametrize("mode", ["fp64_dense", "mixed_fp32"])
def test_energy_dense_exchange_equals_cholesky(c1, mode):
    """The dense exchange block and the Cholesky form give the same E_L."""
    s, p = c1
    nw = p["nwalk"]
    W = perturbed_walkers(s, nw, seed=9)
    out = []
    for vx in (True, False):
        g = m.AfqmcSys(s, nw, mode, dense_exchange=vx)
        g.set(nat.BUF_W, W)
        wd = np.zeros((nw, 8), complex)
        wd[:, 1] = 1
        g.set(nat.BUF_WDATA, wd)
        g.calculate_mixed_density_matrix()
        g.calculate_energy()
        out.append(g.W_data[:, 0])
        g.close()
    close(out[0], out[1], 1e-12 if mode == "fp64_dense" else 1e-5, "E_L dense exchange vs Cholesky")


def run_both(s, p, mode, nsteps=None, nw=None, cutoff=1e-6, csr_dense=True):
    nw = nw or p["nwalk"]
    ns = nsteps or p["nsteps"]
    nsub, northo = p["nsubsteps"], p["northo"]
    x = m.fields_block(p["seed"], s.nchol, nw, 0, ns * nsub)
    g = m.AfqmcSys(s, nw, mode, cutoff=cutoff, csr_dense=csr_dense)
    g.init_walkers()
    g.stage_fields(x)
    g.calculate_mixed_density_matrix()
    e0 = g.calculate_energy()
    eav = g.run(ns, nsub, northo, s.dt)
    res = dict(eav=np.concatenate([[e0], eav]), W=g.W, Wd=g.W_data, eshift=g.eshift())
    g.close()
    spT = orc.halfrotate(s, cutoff=cutoff)
    Lt = orc.dense_Lt(s)
    emode = 0 if (mode == "fp64_csr" and s.vakbl is not None) else 1
    eo, Wo, Wdo, eso, tm, info = orc.run(s, nw, ns, nsub, northo, x, energy_mode=emode, spT=spT, Lt=Lt)
    return res, dict(eav=eo, W=Wo, Wd=Wdo, eshift=eso)


@pytest.mark.parametrize("csr_dense", [True, False])
def test_run_fp64_csr_matches_oracle(c1, csr_dense):
    """The reference's algorithm over 20 steps; the half-rotated SpvnT of this
    system is full, so the bias product runs on its dense copy unless
    AFQMC_OPT_CSR_DENSE is 0 (then as the reference's CSR product)."""
    s, p = c1
    say(f"run fp64_csr N16_M64 20 steps, csr_dense={csr_dense}")
    g, o = run_both(s, p, "fp64_csr", csr_dense=csr_dense)
    close(g["eav"], o["eav"], 1e-9, "Eav trajectory")
    close(g["W"], o["W"], 1e-9, "final walkers")
    close(g["Wd"][:, 0], o["Wd"][:, 0], 1e-9, "final E_L")
    close(g["Wd"][:, 1], o["Wd"][:, 1], 1e-9, "final weights")
    close(g["Wd"][:, 2:4], o["Wd"][:, 2:4], 1e-9, "final overlaps")
    assert abs(g["eshift"] - o["eshift"]) <= 1e-9 * abs(o["eshift"])


def test_run_fp64_dense_matches_oracle(c1):
    s, p = c1
    say("run fp64_dense N16_M64 20 steps")
    g, o = run_both(s, p, "fp64_dense", cutoff=-1.0)
    close(g["eav"], o["eav"], 1e-9, "Eav trajectory")
    close(g["W"], o["W"], 1e-9, "final walkers")
    close(g["Wd"][:, 0], o["Wd"][:, 0], 1e-9, "final E_L")


@pytest.mark.parametrize("mode,tol", [("mixed_fp32", 1e-3), ("mixed_bf16", 1e-3)])
def test_run_mixed_error_vs_fp64_oracle(c1, mode, tol):
    s, p = c1
    say("run", mode, "N16_M64 20 steps")
    g, o = run_both(s, p, mode, cutoff=-1.0)
    dE = np.abs(g["Wd"][:, 0] - o["Wd"][:, 0])
    dEav = np.abs(g["eav"] - o["eav"])
    say(f"{mode}: max |dE_L| = {dE.max():.3e} Ha, max |dEav| = {dEav.max():.3e} Ha")
    assert dE.max() < tol and dEav.max() < tol


def test_energy_terms_sum_and_one_body(c1):
    """E_L splits into one-body + Coulomb + exchange; the one-body term is haj.G (energy.hpp:96-100)."""
    s, p = c1
    nw = p["nwalk"]
    W = perturbed_walkers(s, nw, seed=11)
    for mode in ("fp64_csr", "fp64_dense"):
        g = m.AfqmcSys(s, nw, mode)
        g.set(nat.BUF_W, W)
        wd = np.zeros((nw, 8), complex)
        wd[:, 1] = 1
        g.set(nat.BUF_WDATA, wd)
        G = g.calculate_mixed_density_matrix()
        g.calculate_energy()
        t = g.energy_terms()
        close(t.sum(axis=1), g.W_data[:, 0], 1e-13, f"{mode}: sum of terms vs E_L")
        e1 = np.einsum("iw,i->w", G, s.haj.reshape(-1))
        close(t[:, 0], e1, 1e-12, f"{mode}: one-body term")
        g.close()


def test_coulomb_three_piece(c1):
    """AFQMC_OPT_COULOMB_THREE_PIECE: at energy level 2 of the split-BF16
    mode the Coulomb term's contraction v = Lt.G runs in three-piece
    split-BF16 (24-bit operands).  On identical walkers the one-body and
    exchange terms are the FP64-Coulomb run's bit for bit, and the Coulomb
    term stays within 3e-7 relative (per walker) of the FP64 dense path's
    (measured 9e-8 here, |E_J| about 4 Ha: the 24-bit operands' rounding,
    amplified by the cancellation inside each v_n).
    The option is refused outside the split-BF16 mode and after the upload."""
    s, p = c1
    nw = p["nwalk"]
    W = perturbed_walkers(s, nw, seed=21)
    wd = np.zeros((nw, 8), complex)
    wd[:, 1] = 1
    terms = {}
    for key, mode, c3 in (("fp64", "fp64_dense", False), ("c64", "mixed_bf16", False), ("c3", "mixed_bf16", True)):
        g = m.AfqmcSys(s, nw, mode, energy_precision=2, coulomb_three_piece=c3)
        g.set(nat.BUF_W, W)
        g.set(nat.BUF_WDATA, wd)
        g.calculate_mixed_density_matrix()
        g.profile(True)
        g.calculate_energy()
        sites = set(g.kernel_stats())
        g.profile(False)
        terms[key] = g.energy_terms()
        g.close()
        say(f"{key}: call sites {sorted(x for x in sites if x.startswith('energy_v'))}")
        if mode == "mixed_bf16":
            assert "energy_v_gemm" in sites
    np.testing.assert_array_equal(terms["c3"][:, 0], terms["c64"][:, 0])
    np.testing.assert_array_equal(terms["c3"][:, 2], terms["c64"][:, 2])
    e64 = np.abs(terms["c64"][:, 1] - terms["fp64"][:, 1]) / np.abs(terms["fp64"][:, 1])
    e3 = np.abs(terms["c3"][:, 1] - terms["fp64"][:, 1]) / np.abs(terms["fp64"][:, 1])
    say(f"Coulomb term per-walker rel err: FP64 contraction {e64.max():.2e}, three pieces {e3.max():.2e} "
        f"(|E_J| ~ {np.abs(terms['fp64'][:, 1]).mean():.3f} Ha)")
    assert e64.max() < 1e-12 and e3.max() < 3e-7
    code = {v: k for k, v in nat.STATUS.items()}
    h = C.c_void_p()
    nat.check(nat.lib.afqmc_ctx_create(s.nmo, s.naea, s.nchol, nw, nat.MODES["fp64_dense"], C.byref(h)))
    assert nat.lib.afqmc_set_option(h, nat.OPT_COULOMB_THREE_PIECE, 1.0) == code["AFQMC_ERR_UNSUPPORTED"]
    nat.lib.afqmc_ctx_destroy(h)
    g = m.AfqmcSys(s, nw, "mixed_bf16", energy_precision=2)
    assert nat.lib.afqmc_set_option(g.ctx, nat.OPT_COULOMB_THREE_PIECE, 1.0) == code["AFQMC_ERR_STATE"]
    g.close()


@pytest.mark.parametrize("mode", ["mixed_fp32", "mixed_bf16"])
def test_mixed_energy_evaluation_error(c1, mode):
    """Per-walker E_L of the mixed modes on identical walkers vs the FP64 oracle,
    for each energy precision level (0: all in the contraction precision; 1:
    FP64 G, one-body and Coulomb; 2: and FP32 exchange; 3: and FP64 exchange).
    """
    s, p = c1
    nw = p["nwalk"]
    W = perturbed_walkers(s, nw, seed=13)
    wd = np.zeros((nw, 8), complex)
    wd[:, 1] = 1
    Wdo = wd.copy()
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wdo)
    orc.calculate_energy(s, Wdo, Go)
    eo = Wdo[:, 0]
    for level, tol in ((0, 1e-3), (1, 1e-5), (2, 1e-5), (3, 1e-10)):
        g = m.AfqmcSys(s, nw, mode, energy_precision=level)
        g.set(nat.BUF_W, W)
        g.set(nat.BUF_WDATA, wd)
        g.calculate_mixed_density_matrix()
        g.calculate_energy()
        dE = np.abs(g.W_data[:, 0] - eo)
        say(f"{mode} energy level {level}: max |dE_L| = {dE.max():.3e} Ha (E_L ~ {eo.real.mean():.3f} Ha)")
        assert dE.max() < tol
        g.close()


@pytest.mark.parametrize("mode", ["mixed_fp32", "mixed_bf16"])
def test_exchange_chunks(c1, mode):
    """The dense exchange summed over column chunks of 256 (four chunks of the
    1024 columns here, each chunk's energy added in FP64) against the one-pass
    exchange and the FP64 oracle, for the reduced-precision levels 0-2; chunk
    sizes that are not multiples of 256 are rejected."""
    s, p = c1
    nw = p["nwalk"]
    W = perturbed_walkers(s, nw, seed=29)
    wd = np.zeros((nw, 8), complex)
    wd[:, 1] = 1
    Wdo = wd.copy()
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wdo)
    orc.calculate_energy(s, Wdo, Go)
    eo = Wdo[:, 0]
    for level, tol in ((0, 1e-3), (1, 1e-5), (2, 1e-5)):
        g = m.AfqmcSys(s, nw, mode, energy_precision=level)
        g.set(nat.BUF_W, W)
        g.set(nat.BUF_WDATA, wd)
        g.calculate_mixed_density_matrix()
        e = {}
        for ch in (0, 256, 512):
            g.set_exchange_chunk(ch)
            g.calculate_energy()
            e[ch] = g.W_data[:, 0]
            dE = np.abs(e[ch] - eo)
            say(f"{mode} level {level} chunk {ch}: max |dE_L| = {dE.max():.3e} Ha")
            assert dE.max() < tol
        # the chunks partition the same sum: one-pass and chunked agree to the
        # contraction's own rounding
        assert np.abs(e[256] - e[0]).max() < tol
        assert np.abs(e[512] - e[256]).max() < tol
        with pytest.raises(RuntimeError):
            g.set_exchange_chunk(100)
        g.close()


def test_non_transposed_bias_phases(c1):
    """Full G (mixed_density_matrix.hpp:93-101) and vbias = T(Spvn) G (vbias.hpp:65-82) vs the oracle."""
    s, p = c1
    nw = p["nwalk"]
    g = m.AfqmcSys(s, nw, "fp64_csr", transposed=False)
    W = perturbed_walkers(s, nw, seed=17)
    Wd = np.zeros((nw, 8), complex)
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, Wd)
    say("full G")
    Gf = g.calculate_mixed_density_matrix_full()
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd, compact=False)
    close(Gf, Go, what="full G")
    close(g.W_data[:, 2:4], Wd[:, 2:4], what="overlaps")
    say("vbias from full G")
    vb = g.get_vbias_full()
    vbo = orc.get_vbias(s.spvn, Go, transposed=False)
    close(vb, vbo, what="vbias (non-transposed)")
    g.close()


def test_run_non_transposed_matches_oracle(c1):
    s, p = c1
    nw, ns, nsub, northo = p["nwalk"], 6, 2, 3
    say("run fp64_csr non-transposed N16_M64")
    x = m.fields_block(p["seed"], s.nchol, nw, 0, ns * nsub)
    g = m.AfqmcSys(s, nw, "fp64_csr", transposed=False)
    g.init_walkers()
    g.stage_fields(x)
    g.calculate_mixed_density_matrix()
    e0 = g.calculate_energy()
    eav = g.run(ns, nsub, northo, s.dt)
    W, Wd = g.W, g.W_data
    g.close()
    eo, Wo, Wdo, eso, tm, info = orc.run(s, nw, ns, nsub, northo, x, transposed=False)
    close(np.concatenate([[e0], eav]), eo, 1e-9, "Eav trajectory")
    close(W, Wo, 1e-9, "final walkers")
    close(Wd[:, :4], Wdo[:, :4], 1e-9, "final W_data")


@pytest.mark.parametrize("mode", ["fp64_csr", "fp64_dense", "mixed_bf16"])
def test_graph_replay_matches_eager(c1, mode):
    """Substeps replayed from captured graphs equal the eager substeps, replay after replay,
    with eager work (energy, copies) between the replays and both graph kinds (with ortho)."""
    s, p = c1
    nw, nsub = p["nwalk"], 7
    x = m.fields_block(p["seed"], s.nchol, nw, 0, nsub)
    ctxs = []
    for graphs in (False, True):
        g = m.AfqmcSys(s, nw, mode)
        g.init_walkers()
        g.stage_fields(x)
        if graphs:
            g.use_graphs(True)
        ctxs.append(g)
    say("graph replay", mode)
    for t in range(nsub):
        ortho = t > 0 and t % 3 == 0
        for g in ctxs:
            g.substep(t, ortho, s.dt)
        Wa, Wb = ctxs[0].W, ctxs[1].W
        close(Wb, Wa, 1e-13, f"walkers after substep {t}")
        close(ctxs[1].W_data[:, :5], ctxs[0].W_data[:, :5], 1e-13, f"W_data after substep {t}")
        if t == 4:  # eager work between replays
            for g in ctxs:
                g.calculate_mixed_density_matrix()
                g.calculate_energy()
            close(ctxs[1].W_data[:, 0], ctxs[0].W_data[:, 0], 1e-13, "E_L between replays")
    assert abs(ctxs[1].eshift() - ctxs[0].eshift()) <= 1e-13 * abs(ctxs[0].eshift())
    for g in ctxs:
        g.close()


@pytest.mark.parametrize("nw", [16, 1000])
def test_branch_matches_oracle(c1, nw):
    """Comb population control: parents bit-exact with the oracle's comb, the
    walkers and W_data rows gathered exactly, the common weight W/nw; the
    density matrix of the new walkers then matches the oracle's (the cached
    inverses are rebuilt); degenerate weights are refused with the walkers
    untouched."""
    s, p = c1
    g = m.AfqmcSys(s, nw, "fp64_csr")
    r = np.random.default_rng(nw)
    for u in (0.0, 0.37, 1.0 - 2.0 ** -53):
        W = perturbed_walkers(s, nw, seed=31)
        Wd = r.standard_normal((nw, 8)) + 1j * r.standard_normal((nw, 8))
        Wd[:, 1] = r.exponential(1.0, nw)
        Wd[r.random(nw) < 0.25, 1] = 0.0
        g.set(nat.BUF_W, W)
        g.set(nat.BUF_WDATA, Wd)
        g.calculate_mixed_density_matrix()  # leaves cached inverses of the old walkers
        Wd = g.W_data                        # with the overlaps it wrote
        say(f"branch nw={nw} u={u}")
        parent = g.branch(u)
        Wo, Wdo = W.copy(), Wd.copy()
        st, po = orc.branch(Wo, Wdo, u)
        assert st == 0
        np.testing.assert_array_equal(parent, po)
        np.testing.assert_array_equal(g.W, Wo)
        np.testing.assert_array_equal(g.W_data, Wdo)
    G = g.calculate_mixed_density_matrix()
    Go, _ = orc.calculate_mixed_density_matrix(s, Wo, Wdo.copy())
    close(G, Go, what="G after branching")
    Wd = g.W_data
    Wd[:, 1] = 0.0
    g.set(nat.BUF_WDATA, Wd)
    W0 = g.W
    with pytest.raises(nat.AfqmcError) as ei:
        g.branch(0.5)
    assert ei.value.code == 6
    np.testing.assert_array_equal(g.W, W0)
    g.close()


def test_branch_between_graph_replays(c1):
    """Substeps replayed from graphs, a comb, more replays: equal to the same
    sequence run eagerly (the replay after the comb rebuilds the inverses)."""
    s, p = c1
    nw, nsub = p["nwalk"], 6
    x = m.fields_block(p["seed"], s.nchol, nw, 0, nsub)
    ctxs = []
    for graphs in (False, True):
        g = m.AfqmcSys(s, nw, "fp64_csr")
        g.init_walkers()
        g.stage_fields(x)
        if graphs:
            g.use_graphs(True)
        ctxs.append(g)
    say("branch between graph replays")
    for t in range(nsub):
        for g in ctxs:
            g.substep(t, t == 4, s.dt)
        if t == 2:
            pa, pb = (g.branch(0.61) for g in ctxs)
            np.testing.assert_array_equal(pa, pb)
        close(ctxs[1].W, ctxs[0].W, 1e-13, f"walkers after substep {t}")
        close(ctxs[1].W_data[:, :5], ctxs[0].W_data[:, :5], 1e-13, f"W_data after substep {t}")
    for g in ctxs:
        g.close()


@pytest.mark.parametrize("M,N,K,nw", [(64, 16, 256, 16), (40, 7, 53, 5)])
@pytest.mark.parametrize("mode", ["fp64_csr", "fp64_dense", "mixed_bf16"])
def test_bias_contraction_spins_per_launch(require_gpu, M, N, K, nw, mode):
    """The bench's algorithmic flop count of the bias product assumes one or
    two spin blocks per launch as the library's split-K choice makes them:
    one bias evaluation takes 2 / spins launches of vbias_gemm."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    say("bias launches", M, N, K, nw, mode)
    s = m.make_system(M, N, K, density=0.5, sigma_L=0.03, sigma_h=1.0, dt=0.005, seed=7, with_vakbl=False)
    g = m.AfqmcSys(s, nw, mode, cutoff=1e-6 if mode == "fp64_csr" else -1.0)
    g.profile(True)
    g.get_vbias()
    st = g.kernel_stats()
    g.profile(False)
    g.close()
    spins = 2 if mode == "fp64_csr" else bench.contraction_spins_per_launch(M, N, K, nw)
    f1 = bench.kernel_flops("vbias_gemm", M, N, K, nw, mode)
    say(f"vbias_gemm launches {st.get('vbias_gemm')}, spins per launch {spins}")
    assert st["vbias_gemm"][0] == 2 // spins
    # the flops of one evaluation are both spins' whatever the split
    assert f1 * st["vbias_gemm"][0] == 8.0 * 2 * nw * K * N * M


@pytest.mark.parametrize("M,N,K,nw", [(40, 7, 53, 5), (24, 1, 10, 3), (33, 12, 70, 1)])
@pytest.mark.parametrize("mode", ["fp64_csr", "fp64_dense", "mixed_bf16"])
def test_run_ragged_shapes(require_gpu, M, N, K, nw, mode):
    """Whole steps at shapes that fill no tile evenly (odd M, N, K and walker
    counts, one electron, one walker) against the oracle: FP64 at 1e-9, the
    split-BF16 path within 1e-3 relative on the energies."""
    say("ragged", M, N, K, nw, mode)
    s = m.make_system(M, N, K, density=0.5, sigma_L=0.03, sigma_h=1.0, dt=0.005, seed=M * 100 + N,
                      with_vakbl=(mode == "fp64_csr"))
    p = dict(nsteps=2, nsubsteps=3, northo=2, seed=M + K, nwalk=nw)
    cutoff = 1e-6 if mode == "fp64_csr" else -1.0
    g, o = run_both(s, p, mode, cutoff=cutoff)
    if mode == "mixed_bf16":
        close(g["eav"], o["eav"], 1e-3, "Eav trajectory")
    else:
        close(g["eav"], o["eav"], 1e-9, "Eav trajectory")
        close(g["W"], o["W"], 1e-9, "final walkers")
        close(g["Wd"][:, :5], o["Wd"][:, :5], 1e-9, "final W_data")


def test_ragged_split_bf16_energy_level2(require_gpu):
    """Energy level 2 of the split-BF16 mode at a shape whose N*M (280) is
    no multiple of 16: the exchange takes the FP32 block (the three-piece
    one needs 16-complex k-tiles) and E_L stays within 1e-5 Ha of the
    oracle; the three-piece contraction options are refused at such shapes
    (M = 40 for the propagator, K = 53 for the HS potential, N*M = 280 for
    the bias)."""
    M, N, K, nw = 40, 7, 53, 5
    say("ragged split-BF16 energy level 2", M, N, K, nw)
    s = m.make_system(M, N, K, density=0.5, sigma_L=0.03, sigma_h=1.0, dt=0.005, seed=M * 100 + N, with_vakbl=True)
    W = perturbed_walkers(s, nw, seed=5)
    wd = np.zeros((nw, 8), complex)
    wd[:, 1] = 1
    Wdo = wd.copy()
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wdo)
    orc.calculate_energy(s, Wdo, Go)
    g = m.AfqmcSys(s, nw, "mixed_bf16", energy_precision=2)
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, wd)
    g.calculate_mixed_density_matrix()
    g.calculate_energy()
    dE = np.abs(g.W_data[:, 0] - Wdo[:, 0])
    say(f"  max |dE_L| = {dE.max():.3e} Ha (E_L ~ {Wdo[:, 0].real.mean():.3f} Ha)")
    assert dE.max() < 1e-5
    g.close()
    code = {v: k for k, v in nat.STATUS.items()}
    h = C.c_void_p()
    nat.check(nat.lib.afqmc_ctx_create(M, N, K, nw, nat.MODES["mixed_bf16"], C.byref(h)))
    for mask in (4, 2, 1):
        assert nat.lib.afqmc_set_option(h, nat.OPT_THREE_PIECE, float(mask)) == code["AFQMC_ERR_UNSUPPORTED"]
    nat.lib.afqmc_ctx_destroy(h)


@pytest.mark.parametrize("bad", ["column", "rowptr"])
def test_upload_rejects_malformed_spvn(require_gpu, bad):
    """A Cholesky CSR with an out-of-range column or a decreasing row pointer
    is refused at upload (AFQMC_ERR_SHAPE) before any kernel reads it."""
    s = m.make_system(12, 3, 8, density=0.5, sigma_L=0.03, seed=5)
    col, rp = s.spvn.col.copy(), s.spvn.rowptr.copy()
    if bad == "column":
        col[len(col) // 2] = s.nchol
    else:
        rp[5] = rp[6] + 1  # row 5 would end before it starts
    s.spvn = m.Csr(s.spvn.nrows, s.spvn.ncols, rp, col, s.spvn.val)
    with pytest.raises(nat.AfqmcError) as ei:
        m.AfqmcSys(s, 2, "fp64_csr")
    assert ei.value.code == 2


@pytest.mark.parametrize("shape", ["N16_M64", "N64_M256", "N128_M512"])
def test_fused_propagator_matches_gemm_chain(require_gpu, shape):
    """The one-kernel split-BF16 propagator gives the walkers of the eight-GEMM
    chain bit for bit (same operand splits, MFMA order and FP64 epilogues)."""
    s, p = m.preset_system(shape)
    nw = 16
    W = perturbed_walkers(s, nw, seed=41, noise=0.05)
    x = m.fields(p["seed"], s.nchol, nw, 0)
    out = []
    for fused in (False, True):
        say(f"propagator {shape} fused={fused}")
        g = m.AfqmcSys(s, nw, "mixed_bf16", dense_exchange=False)
        g.set_fused_propagator(fused)
        g.set(nat.BUF_W, W)
        g.calculate_mixed_density_matrix()
        g.get_vbias()
        g.fields(x)
        g.get_vHS()
        g.propagate()
        out.append(g.W)
        g.close()
    assert np.array_equal(out[0], out[1]), f"max diff {np.abs(out[0] - out[1]).max():.3e}"


@pytest.mark.parametrize("shape", [(40, 8, 64), (48, 12, 96), "N64_M256", "N128_M512"])
def test_direct_lds_staging_matches_register_path(require_gpu, shape):
    """Split-BF16 GEMMs whose operands go from global memory straight into
    LDS (swizzled tiles; level 1 two stages for the left operand, level 2 a
    ring of stages for both) give the walkers, overlaps and energies of the
    register-staged path bit for bit; 40 and 48 orbitals leave partial row
    tiles (rows past m read a clamped row whose outputs are dropped), and
    at 40 orbitals the propagator's k = 40 is not a multiple of 16, so only
    the bias and vHS products take the direct path."""
    if isinstance(shape, str):
        s, p = m.preset_system(shape)
        seed = p["seed"]
    else:
        M, N, K = shape
        s = m.make_system(M, N, K, density=0.5, sigma_L=0.02, seed=9)
        seed = 9
    nw = 16
    x = m.fields_block(seed, s.nchol, nw, 0, 4)
    out = []
    for direct in (0, 1, 2):
        say(f"direct LDS staging {shape} level={direct}")
        g = m.AfqmcSys(s, nw, "mixed_bf16", dense_exchange=False)
        g.set_direct_lds(direct)
        g.init_walkers()
        g.stage_fields(x)
        g.calculate_mixed_density_matrix()
        g.calculate_energy()
        g.run(1, 4, 2, s.dt)
        out.append((g.W, g.W_data))
        g.close()
    for lv in (1, 2):
        for a, b in zip(out[0], out[lv]):
            assert np.array_equal(a, b), f"level {lv}: max diff {np.abs(a - b).max():.3e}"


@pytest.mark.parametrize("tile", [1, 2, 3, 4])
def test_every_ring_tile_matches_register_path(require_gpu, tile):
    """Every tile shape of the staged-ring split-BF16 GEMM (forced on all call
    sites) gives the walkers, overlaps and local energies of the
    register-staged path bit for bit at 512 orbitals / 128 electrons (the
    Taylor steps, P U, vHS, the bias product and the triangular exchange
    block in split-BF16 at energy level 0, exchange chunks of 2048)."""
    s = m.make_system(512, 128, 256, density=1.0, sigma_L=0.02 * (64.0 * 256 / (512 * 256)) ** 0.5, seed=31)
    nw = 24
    x = m.fields_block(31, s.nchol, nw, 0, 3)
    out = []
    for direct in (0, 2):
        say(f"ring tile {tile}, level {direct}")
        g = m.AfqmcSys(s, nw, "mixed_bf16", energy_precision=0)
        g.set_exchange_chunk(2048)
        g.set_direct_lds(direct)
        g.set_tile(tile)
        g.init_walkers()
        g.stage_fields(x)
        g.calculate_mixed_density_matrix()
        g.calculate_energy()
        g.run(1, 3, 2, s.dt)
        out.append((g.W, g.W_data))
        g.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b), f"tile {tile}: max diff {np.abs(a - b).max():.3e}"


def test_repeated_spvn_entries_are_summed(require_gpu):
    """A Spvn with repeated (row, column) entries (the reference keeps them:
    SparseMatrix::compress only sorts, SparseMatrix.hpp:295-319, and csrmm
    sums them, sparse.hpp:166-180) gives the walkers and energies of the same
    matrix with each repeated entry summed, in every mode: every entry of the
    first 40 rows is split into two halves, and row 7 repeats its first entry
    with value zero."""
    import dataclasses
    from miniafqmc_amd.system import Csr
    s = m.make_system(24, 6, 40, density=0.5, sigma_L=0.03, sigma_h=1.0, dt=0.005, seed=77, with_vakbl=True)
    sp = s.spvn
    rows, cols, vals = [], [], []
    for r in range(sp.nrows):
        for q in range(sp.rowptr[r], sp.rowptr[r + 1]):
            if r < 40:
                rows += [r, r]
                cols += [sp.col[q], sp.col[q]]
                vals += [sp.val[q] / 2, sp.val[q] / 2]
            else:
                rows.append(r)
                cols.append(sp.col[q])
                vals.append(sp.val[q])
            if r == 7 and q == sp.rowptr[r]:
                rows.append(r)
                cols.append(sp.col[q])
                vals.append(0.0)
    rp = np.zeros(sp.nrows + 1, np.int32)
    np.add.at(rp, np.asarray(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    dup = Csr(sp.nrows, sp.ncols, rp, np.asarray(cols, np.int32), np.asarray(vals, np.complex128))
    assert np.array_equal(dup.dense(), sp.dense())
    s2 = dataclasses.replace(s, spvn=dup)
    nw, nsub = 6, 3
    x = m.fields_block(77, s.nchol, nw, 0, nsub)
    for mode, tol in (("fp64_csr", 1e-12), ("fp64_dense", 0.0), ("mixed_bf16", 0.0)):
        out = []
        for sysm in (s, s2):
            g = m.AfqmcSys(sysm, nw, mode)
            g.init_walkers()
            g.stage_fields(x)
            g.calculate_mixed_density_matrix()
            g.run(1, nsub, 2, s.dt)
            out.append((g.W, g.W_data))
            g.close()
        for a, b in zip(out[0], out[1]):
            d = np.abs(a - b).max() / np.abs(a).max()
            say(f"repeated Spvn entries, {mode}: max rel diff {d:.2e}")
            assert d <= tol


@pytest.mark.parametrize("mode", ["fp64_csr", "fp64_dense", "mixed_bf16"])
def test_log_overlaps_keep_ratios_beyond_fp64_range(c1, mode):
    """The substep is invariant under scaling the walkers by a constant: G,
    the bias, vHS and the propagation are, and the weights take the overlap
    ratio from log-determinants.  Walkers scaled by 2^-73 (about 1e-22; a
    power of two, so every product scales without rounding) have overlap
    determinants of about 1e-352 per spin, zero as FP64 products, and must
    give the same weights, hybrid energies, Eshift and orthogonalised
    walkers as the unscaled ones; what differs is the rounding of the
    log-determinant sums (about 1e3 each here)."""
    s, p = c1
    nw, nsub = 8, 4
    W = perturbed_walkers(s, nw, seed=7, noise=0.05)
    x = m.fields_block(p["seed"], s.nchol, nw, 0, nsub)
    out = []
    for scale in (1.0, 2.0 ** -73):
        say(f"log overlaps {mode} scale {scale:g}")
        g = m.AfqmcSys(s, nw, mode)
        g.stage_fields(x)
        Wd = np.zeros((nw, 8), complex)
        Wd[:, 1] = 1.0
        g.set(nat.BUF_W, W * scale)
        g.set(nat.BUF_WDATA, Wd)
        g.calculate_mixed_density_matrix()
        raw = g.W_data[:, 2:4].copy()
        for t in range(nsub):
            g.substep(t, t == nsub - 1, s.dt)
        out.append((raw, g.W_data, g.W, g.eshift()))
        g.close()
    (raw1, wd1, w1, es1), (raws, wds, ws, ess) = out
    assert np.all(np.abs(raw1) > 0)
    assert np.all(raws == 0), "the scaled overlaps should underflow as FP64 products"
    assert np.all(np.isfinite(wds[:, [1, 4, 5]]))
    tol = 1e-10
    close(wds[:, 1], wd1[:, 1], tol, "weights")
    close(wds[:, 4:6], wd1[:, 4:6], tol, "hybrid energies")
    close(ws, w1, tol, "orthogonalised walkers")
    assert abs(ess - es1) <= tol * abs(es1)


@pytest.mark.parametrize("mode", ["fp64_csr", "fp64_dense", "mixed_bf16"])
@pytest.mark.parametrize("graphs", [False, True])
def test_device_fields_equal_host_draws(c1, mode, graphs):
    """Fields drawn on the device inside the substep (afqmc_device_fields)
    are the host generator's draws bit for bit: the same X, hybrid weights
    and walkers as fields staged from the host (miniafqmc_base.cpp:282)."""
    s, p = c1
    nw, nsub = 12, 3
    out = []
    for device in (False, True):
        say(f"device fields {mode} graphs={graphs} device={device}")
        g = m.AfqmcSys(s, nw, mode)
        if device:
            g.device_fields(p["seed"])
        else:
            g.stage_fields(m.fields_block(p["seed"], s.nchol, nw, 0, nsub))
        g.init_walkers()
        g.calculate_mixed_density_matrix()
        if graphs:
            g.use_graphs(True)
        for t in range(nsub):
            g.substep(t, t == 2, s.dt)
        out.append((g.get(nat.BUF_X), g.get(nat.BUF_HYBRIDW), g.W, g.W_data))
        g.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_csr_dense_bias_matches_csr_product(c1):
    """AFQMC_OPT_CSR_DENSE: with SpvnT at least half full the FP64_CSR bias
    product runs on a dense copy (entries under the cutoff zero), split into
    chunks summed in FP64; it gives the CSR product's vbias to 1e-13, and a
    sparser SpvnT (a cutoff that drops most entries) keeps the CSR product."""
    s, p = c1
    nw = p["nwalk"]
    W = perturbed_walkers(s, nw, seed=13)
    Wd = np.zeros((nw, 8), complex)
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
    out = {}
    for dense in (True, False):
        g = m.AfqmcSys(s, nw, "fp64_csr", csr_dense=dense)
        g.set(nat.BUF_G, Go)
        out[dense] = g.get_vbias()
        g.close()
    close(out[True], out[False], 1e-13, "dense vs CSR bias product")
    close(out[True], orc.get_vbias(orc.halfrotate(s), Go), RTOL, "dense bias product vs oracle")
    spT = orc.halfrotate(s, cutoff=1e-6)
    say(f"SpvnT density at cutoff 1e-6: {spT.nnz / (spT.nrows * spT.ncols):.3f}")
    # which product ran: the call sites' kernel statistics
    full = orc.halfrotate(s, cutoff=-1.0)
    cut = float(np.quantile(np.abs(full.val), 0.7))  # keeps about 30 % of the entries
    for cutoff, site in ((1e-6, "vbias_gemm"), (cut, "csrmm_vbias")):
        g = m.AfqmcSys(s, nw, "fp64_csr", cutoff=cutoff)
        g.set(nat.BUF_G, Go)
        g.profile(True)
        vb = g.get_vbias()
        names = set(g.kernel_stats())
        g.profile(False)
        g.close()
        say(f"cutoff {cutoff:.3g}: call sites {sorted(names)}")
        assert site in names
        close(vb, orc.get_vbias(orc.halfrotate(s, cutoff=cutoff), Go), RTOL, f"bias product at cutoff {cutoff:.3g}")
