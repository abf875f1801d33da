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
