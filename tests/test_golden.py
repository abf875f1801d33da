"""Committed fixture tests/golden/walker_step_m16.npz (made by
tests/golden/make_golden.py): one walker step of a 16-orbital system, inputs
and outputs phase by phase, and a 6-substep trajectory.

CPU: the oracle still reproduces the fixture and the synthetic generator
still reproduces its inputs.  GPU: the HIP path, fed the fixture's inputs
through the C-ABI, reproduces its outputs (FP64 at 1e-10, trajectory 1e-9).
"""
import os

import numpy as np
import pytest

import miniafqmc_amd as m
import oracle_lib as orc
from miniafqmc_amd import _native as nat
from miniafqmc_amd.system import Csr, System

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walker_step_m16.npz")


@pytest.fixture(scope="module")
def gold():
    z = np.load(PATH)
    M, N, K = int(z["nmo"]), int(z["naea"]), int(z["nchol"])
    s = System(M, N, K, float(z["dt"]), int(z["seed"]), z["trial_alpha"], z["trial_beta"], z["haj"], z["propg1"],
               z["h1"], Csr(M * M, K, z["spvn_rowptr"], z["spvn_col"], z["spvn_val"]),
               Csr(2 * N * M, 2 * N * M, z["vakbl_rowptr"], z["vakbl_col"], z["vakbl_val"]))
    return s, z


def close(a, b, rtol, what):
    a, b = np.asarray(a), np.asarray(b)
    err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
    assert err <= rtol, f"{what}: rel err {err:.3e} > {rtol:.1e}"


def test_generator_reproduces_inputs(gold):
    s, z = gold
    g = m.make_system(s.nmo, s.naea, s.nchol, density=0.5, sigma_L=0.05, sigma_h=1.0, dt=s.dt, seed=s.seed,
                      with_vakbl=True)
    close(g.trial_alpha, s.trial_alpha, 1e-14, "trial_alpha")
    close(g.propg1, s.propg1, 1e-14, "Propg1")
    assert np.array_equal(g.spvn.col, s.spvn.col) and np.array_equal(g.spvn.rowptr, s.spvn.rowptr)
    close(g.spvn.val, s.spvn.val, 1e-14, "Spvn")
    close(g.vakbl.val, s.vakbl.val, 1e-14, "Vakbl")
    assert np.array_equal(m.fields(s.seed, s.nchol, 4, 0), z["x"])


def test_oracle_reproduces_fixture(gold):
    s, z = gold
    W, Wd = z["W"].copy(), z["Wd"].copy()
    spT = orc.halfrotate(s, cutoff=1e-6)
    assert np.array_equal(spT.col, z["spvnT_col"])
    close(spT.val, z["spvnT_val"], 1e-13, "SpvnT")
    G, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
    close(G, z["G"], 1e-12, "G")
    close(Wd[:, 2:4], z["ovlp"], 1e-12, "overlaps")
    vb = orc.get_vbias(spT, G)
    close(vb, z["vbias"], 1e-12, "vbias")
    X, hw = orc.fields(vb, z["x"])
    close(X, z["X"], 1e-12, "X")
    vhs = orc.get_vhs(s.spvn, X)
    close(vhs, z["vHS"], 1e-12, "vHS")
    Wp = orc.propagate(s, W, vhs)
    close(Wp, z["W_prop"], 1e-12, "propagated W")
    close(orc.orthogonalize(Wp), z["W_ortho"], 1e-12, "orthogonalized W")


@pytest.mark.gpu
def test_gpu_phases_match_fixture(gold, require_gpu):
    s, z = gold
    nw = z["W"].shape[0]
    g = m.AfqmcSys(s, nw, "fp64_csr")
    g.set(nat.BUF_W, z["W"])
    g.set(nat.BUF_WDATA, z["Wd"])
    close(g.calculate_mixed_density_matrix(), z["G"], 1e-10, "G")
    close(g.W_data[:, 2:4], z["ovlp"], 1e-10, "overlaps")
    close(g.get_vbias(), z["vbias"], 1e-10, "vbias")
    g.fields(z["x"])
    close(g.get(nat.BUF_X), z["X"], 1e-10, "X")
    close(g.get(nat.BUF_HYBRIDW), z["hybridW"], 1e-10, "hybridW")
    close(g.get_vHS(), z["vHS"], 1e-10, "vHS")
    g.propagate()
    close(g.W, z["W_prop"], 1e-10, "propagated W")
    g.save_overlaps()
    g.calculate_overlaps()
    es = g.weights(s.dt, 0.0)
    close(g.W_data, z["Wd_weights"], 1e-10, "W_data after weights")
    assert abs(es - float(z["eshift"])) <= 1e-10 * abs(float(z["eshift"]))
    g.orthogonalize()
    close(g.W, z["W_ortho"], 1e-10, "orthogonalized W")
    g.calculate_mixed_density_matrix()
    close(g.get(nat.BUF_G), z["G_ortho"], 1e-10, "G after ortho")
    eav = g.calculate_energy()
    close(g.W_data[:, 0], z["E_L"], 1e-10, "E_L")
    assert abs(eav - float(z["Eav"])) <= 1e-10 * abs(float(z["Eav"]))
    g.close()


@pytest.mark.gpu
def test_gpu_trajectory_matches_fixture(gold, require_gpu):
    s, z = gold
    nw = z["W"].shape[0]
    g = m.AfqmcSys(s, nw, "fp64_csr")
    g.init_walkers()
    g.stage_fields(z["run_x"])
    g.calculate_mixed_density_matrix()
    e0 = g.calculate_energy()
    eav = g.run(1, int(z["run_nsub"]), int(z["run_northo"]), s.dt)
    tol = 1e-9
    close(np.concatenate([[e0], eav]), z["run_eav"], tol, "Eav")
    close(g.W, z["run_W"], tol, "final walkers")
    close(g.W_data[:, :4], z["run_Wd"][:, :4], tol, "final W_data")
    g.close()
