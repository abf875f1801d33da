"""Three-piece split-BF16 contractions (AFQMC_OPT_THREE_PIECE, split-BF16 mode):
the bias product, the HS potential and the propagator with 24-bit operands
(x = h + m + l, six bf16 products accumulated in FP32) against the FP64
oracle, phase by phase from shared inputs and over a 20-step walk; the
bounds are written below and sit between the FP64 path (1e-10) and the
two-piece split (about 1e-5 per phase).  Phases follow
src/AFQMC/vbias.hpp:45-83, vHS.hpp:43-91 and afqmc_sys.hpp:146-173.
"""
import sys

import numpy as np
import pytest

import miniafqmc_amd as m
import oracle_lib as orc
from miniafqmc_amd import _native as nat

pytestmark = pytest.mark.gpu

# relative bounds of the three-piece phases (max |GPU - oracle| / max |oracle|;
# measured 9e-9 (X) to 7e-7 (vHS, W), against 3e-6 to 2e-5 for two pieces)
PHASE_BOUND = 1e-6
# 20-step walk of 16 walkers at 64 orbitals: per-walker local energy and Eav
# (measured 1.5e-6 / 2.7e-7 Ha, against 9.4e-5 / 1.1e-5 Ha for two pieces)
WALK_DE_L = 5e-6  # Ha


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def c1(require_gpu):
    return m.preset_system("N16_M64")


def perturbed_walkers(s, nw, seed=3, noise=0.1):
    r = np.random.default_rng(seed)
    W = np.zeros((nw, 2, s.nmo, s.naea), complex)
    W[:, 0] = s.trial_alpha.conj()
    W[:, 1] = s.trial_beta.conj()
    W += noise * (r.standard_normal(W.shape) + 1j * r.standard_normal(W.shape))
    return W


def phase_errors(s, p, three):
    nw = p["nwalk"]
    g = m.AfqmcSys(s, nw, "mixed_bf16", three_piece=three)
    spT = orc.halfrotate(s, cutoff=-1.0)
    W = perturbed_walkers(s, nw, seed=5)
    Wd = np.zeros((nw, 8), complex)
    Wd[:, 1] = 1.0
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, Wd)
    err = {}
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd.copy())
    err["G"] = rel(g.calculate_mixed_density_matrix(), Go)
    g.set(nat.BUF_G, Go)
    vbo = orc.get_vbias(spT, Go)
    err["vbias"] = rel(g.get_vbias(), vbo)
    x = m.fields(p["seed"], s.nchol, nw, 0)
    g.set(nat.BUF_VBIAS, vbo)
    g.fields(x)
    Xo, ho = orc.fields(vbo, x)
    err["X"] = rel(g.get(nat.BUF_X), Xo)
    g.set(nat.BUF_X, Xo)
    vhso = orc.get_vhs(s.spvn, Xo)
    err["vHS"] = rel(g.get_vHS(), vhso)
    g.set(nat.BUF_VHS, vhso)
    g.propagate()
    err["W"] = rel(g.W, orc.propagate(s, W, vhso))
    g.close()
    return err


def test_three_piece_phases_vs_oracle(c1):
    s, p = c1
    two = phase_errors(s, p, 0)
    three = phase_errors(s, p, 7)
    for k in three:
        say(f"{k}: two-piece {two[k]:.2e}, three-piece {three[k]:.2e}")
        assert three[k] < PHASE_BOUND, (k, three[k])
    # every product phase gains at least a factor of five
    for k in ("vbias", "vHS", "W"):
        assert three[k] < 0.2 * two[k], (k, two[k], three[k])


def test_three_piece_walk_vs_oracle(c1):
    s, p = c1
    nw, ns, nsub, northo = p["nwalk"], p["nsteps"], p["nsubsteps"], p["northo"]
    x = m.fields_block(p["seed"], s.nchol, nw, 0, ns * nsub)
    out = {}
    for three in (0, 7):
        g = m.AfqmcSys(s, nw, "mixed_bf16", energy_precision=3, three_piece=three)
        g.init_walkers()
        g.stage_fields(x)
        g.calculate_mixed_density_matrix()
        e0 = g.calculate_energy()
        eav = g.run(ns, nsub, northo, s.dt)
        out[three] = (np.concatenate([[e0], eav]), g.W_data[:, 0])
        g.close()
    spT = orc.halfrotate(s, cutoff=-1.0)
    eo, Wo, Wdo, eso, tm, info = orc.run(s, nw, ns, nsub, northo, x, energy_mode=1, spT=spT, Lt=orc.dense_Lt(s))
    for three, (eav, el) in out.items():
        dE, dEav = np.abs(el - Wdo[:, 0]).max(), np.abs(eav - eo).max()
        say(f"three-piece mask {three}: 20 steps, max |dE_L| {dE:.2e} Ha, max |dEav| {dEav:.2e} Ha")
        if three:
            assert dE < WALK_DE_L and dEav < WALK_DE_L


def test_three_piece_graph_replay_and_tiles(c1):
    """Captured substeps replay the three-piece kernels like eager launches, and
    every staged-ring tile shape gives the same walkers bit for bit."""
    s, p = c1
    nw = p["nwalk"]
    x = m.fields_block(p["seed"], s.nchol, nw, 0, 6)
    out = []
    for graphs, tile in ((False, 0), (True, 0), (False, 1), (False, 2), (False, 3), (False, 4)):
        say(f"three-piece graphs={graphs} tile={tile}")
        g = m.AfqmcSys(s, nw, "mixed_bf16", three_piece=7)
        g.set_tile(tile)
        g.init_walkers()
        g.stage_fields(x)
        g.calculate_mixed_density_matrix()
        if graphs:
            g.use_graphs(True)
        g.run(2, 3, 2, s.dt, energy=False)
        out.append((g.W, g.W_data))
        g.close()
    for i, (a, b) in enumerate(out[1:]):
        assert np.array_equal(a, out[0][0]) and np.array_equal(b, out[0][1]), f"variant {i + 1} differs"


def test_three_piece_option_checks(c1):
    s, p = c1
    g = m.AfqmcSys(s, 4, "fp64_dense")
    with pytest.raises(nat.AfqmcError):
        nat.check(nat.lib.afqmc_set_option(g.ctx, nat.OPT_THREE_PIECE, 7.0), g.ctx)
    g.close()


def propagate_error(s, p, mask):
    """Relative error of the propagated walkers against the oracle, with the
    propagator in three pieces on the products `mask` selects (bit 0 P W,
    bits 1-6 the Horner steps U_6 .. U_1, bit 7 P U)."""
    nw = p["nwalk"]
    g = m.AfqmcSys(s, nw, "mixed_bf16", three_piece=4)
    g.set_three_piece_steps(mask)
    W = perturbed_walkers(s, nw, seed=11)
    spT = orc.halfrotate(s, cutoff=-1.0)
    Wd = np.zeros((nw, 8), complex)
    Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
    Xo, _ = orc.fields(orc.get_vbias(spT, Go), m.fields(p["seed"], s.nchol, nw, 0))
    vhso = orc.get_vhs(s.spvn, Xo)
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_VHS, vhso)
    g.propagate()
    e = rel(g.W, orc.propagate(s, W, vhso))
    g.close()
    return e


def test_three_piece_propagator_step_masks(c1):
    """The propagator's products take three pieces where the step mask says:
    all eight (255, the default) meet the three-piece bound, none (0) sits at
    the two-piece level, and P W and P U alone (0x81) lie between."""
    s, p = c1
    e = {mk: propagate_error(s, p, mk) for mk in (0x00, 0x81, 0xC1, 0xFF)}
    say("propagator step masks: " + ", ".join(f"0x{k:02x} {v:.2e}" for k, v in e.items()))
    assert e[0xFF] < PHASE_BOUND
    assert e[0x00] > 5 * e[0xFF]
    assert e[0xFF] <= e[0xC1] <= e[0x81] * 1.05 and e[0x81] < e[0x00]


def test_three_piece_step_mask_option(c1):
    """Mask 255 is the default bit for bit; a mask selects launches that
    graph replays follow; masks beyond eight bits are refused."""
    s, p = c1
    nw = p["nwalk"]
    x = m.fields_block(p["seed"], s.nchol, nw, 0, 6)
    out = []
    for mask, graphs in ((None, False), (0xFF, False), (0x81, False), (0x81, True)):
        g = m.AfqmcSys(s, nw, "mixed_bf16", three_piece=4)
        if mask is not None:
            g.set_three_piece_steps(mask)
        g.init_walkers()
        g.stage_fields(x)
        g.calculate_mixed_density_matrix()
        if graphs:
            g.use_graphs(True)
        g.run(2, 3, 2, s.dt, energy=False)
        out.append(g.W)
        if mask == 0x81 and graphs:
            with pytest.raises(nat.AfqmcError):
                nat.check(nat.lib.afqmc_set_option(g.ctx, nat.OPT_THREE_PIECE_STEPS, 256.0), g.ctx)
        g.close()
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[1], out[2])
    assert np.array_equal(out[2], out[3])


def test_one_piece_horner_steps(c1):
    """The innermost Horner steps in one piece (the hi bf16 words): their
    terms carry V^k/k!, so the propagated walkers stay at the two-piece
    level against the oracle, while all six steps in one piece do not;
    graph replays follow the mask, and P W / P U cannot take one piece."""
    s, p = c1
    e = {}
    for one in (0x00, 0x0E, 0x7E):
        nw = p["nwalk"]
        g = m.AfqmcSys(s, nw, "mixed_bf16", three_piece=4)
        g.set_three_piece_steps(0x81)
        g.set_one_piece_steps(one)
        W = perturbed_walkers(s, nw, seed=11)
        spT = orc.halfrotate(s, cutoff=-1.0)
        Wd = np.zeros((nw, 8), complex)
        Go, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
        Xo, _ = orc.fields(orc.get_vbias(spT, Go), m.fields(p["seed"], s.nchol, nw, 0))
        vhso = orc.get_vhs(s.spvn, Xo)
        g.set(nat.BUF_W, W)
        g.set(nat.BUF_VHS, vhso)
        g.propagate()
        e[one] = rel(g.W, orc.propagate(s, W, vhso))
        if one == 0x0E:
            for bad in (0x01, 0x80, 256):
                with pytest.raises(nat.AfqmcError):
                    nat.check(nat.lib.afqmc_set_option(g.ctx, nat.OPT_ONE_PIECE_STEPS, float(bad)), g.ctx)
        g.close()
    say("one-piece Horner steps: " + ", ".join(f"0x{k:02x} {v:.2e}" for k, v in e.items()))
    assert e[0x0E] < 2 * e[0x00] + 1e-7
    assert e[0x7E] > 5 * e[0x0E]
    nw = p["nwalk"]
    x = m.fields_block(p["seed"], s.nchol, nw, 0, 6)
    out = []
    for graphs in (False, True):
        g = m.AfqmcSys(s, nw, "mixed_bf16", three_piece=4)
        g.set_three_piece_steps(0x81)
        g.set_one_piece_steps(0x0E)
        g.init_walkers()
        g.stage_fields(x)
        g.calculate_mixed_density_matrix()
        if graphs:
            g.use_graphs(True)
        g.run(2, 3, 2, s.dt, energy=False)
        out.append(g.W)
        g.close()
    assert np.array_equal(out[0], out[1])
