"""Declared error paths through the C-ABI, against the oracle's status.

The reference only asserts (LAPACK info > 0 behind ma::invert and
ma::determinant, src/Numerics/ma_lapack.hpp:45,77; NaN draws pass through its
fields loop, src/miniapps/miniafqmc_base.cpp:281-289).  The library turns both
into status codes read from device flags after the kernels have ended
normally, and the oracle reports the same two conditions:
  - an exactly singular overlap matrix (a walker with a zero column):
    AFQMC_ERR_SINGULAR on the GPU, LAPACK info = the zero pivot's 1-based
    column in the oracle; the other walkers' results are unaffected;
  - a non-finite auxiliary-field draw: AFQMC_ERR_NONFINITE on the GPU,
    ORC_NONFINITE from the oracle's driver loop;
  - a rank-deficient walker through the orthogonalisation has no error code in
    either; the LQ factor's columns before the dependent one agree and the
    result is orthonormal.
The flags are cleared after each report, so the next call succeeds.
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


@pytest.fixture(scope="module")
def c1(require_gpu):
    return m.preset_system("N16_M64")


def walkers(s, nw, seed=3, noise=0.1):
    r = np.random.default_rng(seed)
    W = np.zeros((nw, 2, s.nmo, s.naea), complex)
    W[:, 0] = s.trial_alpha.conj()
    W[:, 1] = s.trial_beta.conj()
    W += noise * (r.standard_normal(W.shape) + 1j * r.standard_normal(W.shape))
    return W


@pytest.mark.parametrize("mode", ["fp64_csr", "fp64_dense", "mixed_bf16"])
def test_singular_overlap(c1, mode):
    s, p = c1
    nw = 4
    say(f"singular overlap matrix, {mode}")
    W = walkers(s, nw)
    W[2, 0, :, 3] = 0.0  # walker 2, spin up: column 3 zero -> T1 has a zero row
    Wd = np.zeros((nw, 8), complex)
    Wd[:, 1] = 1.0
    Go, info = orc.calculate_mixed_density_matrix(s, W, Wd.copy())
    assert info == 4  # zgetf2: first zero pivot in column 4 (1-based)
    g = m.AfqmcSys(s, nw, mode)
    g.set(nat.BUF_W, W)
    g.set(nat.BUF_WDATA, Wd)
    with pytest.raises(nat.AfqmcError) as e:
        g.calculate_mixed_density_matrix()
    assert e.value.code == 5  # AFQMC_ERR_SINGULAR
    G = g.get(nat.BUF_G)
    ok = [0, 1, 3]
    tol = 1e-10 if mode.startswith("fp64") else 1e-4
    assert rel(G[:, ok], Go[:, ok]) < tol
    assert g.W_data[2, 2] == 0  # det of the singular matrix
    with pytest.raises(nat.AfqmcError) as e:
        g.calculate_overlaps()
    assert e.value.code == 5
    ovo = np.zeros((nw, 8), complex)
    assert orc.calculate_overlaps(s, W, ovo) == 4
    assert ovo[2, 2] == 0
    assert rel(g.W_data[ok, 2:4], ovo[ok, 2:4]) < 1e-10
    # the flag is cleared: regular walkers go through
    W[2, 0, :, 3] = W[1, 0, :, 3]
    g.set(nat.BUF_W, W)
    g.calculate_mixed_density_matrix()
    g.close()


@pytest.mark.parametrize("mode", ["fp64_csr", "mixed_bf16"])
@pytest.mark.parametrize("graphs", [False, True])
def test_nonfinite_fields(c1, mode, graphs):
    s, p = c1
    nw, nsub = 4, 3
    say(f"non-finite field draw, {mode}, graphs={graphs}")
    x = m.fields_block(p["seed"], s.nchol, nw, 0, nsub)
    x[1, 5, 2] = np.nan  # substep 1, vector 5, walker 2
    g = m.AfqmcSys(s, nw, mode)
    g.init_walkers()
    g.stage_fields(x)
    if graphs:
        g.use_graphs(True)
    with pytest.raises(nat.AfqmcError) as e:
        g.run(1, nsub, 10, s.dt, energy=False)
    assert e.value.code == 6  # AFQMC_ERR_NONFINITE
    eo, Wo, Wdo, _, _, info = orc.run(s, nw, 1, nsub, 10, x, energy_mode=2)
    assert info == orc.ORC_NONFINITE
    # only walker 2 is poisoned; the others are the oracle's
    W = g.W
    assert not np.all(np.isfinite(W[2]))
    ok = [0, 1, 3]
    assert rel(W[ok], Wo[ok]) < (1e-9 if mode == "fp64_csr" else 1e-3)
    # single fields call: reported, then cleared
    g.init_walkers()
    g.calculate_mixed_density_matrix()
    g.get_vbias()
    with pytest.raises(nat.AfqmcError) as e:
        g.fields(x[1])
    assert e.value.code == 6
    g.fields(x[0])
    g.close()


def test_rank_deficient_orthogonalize(c1):
    s, p = c1
    nw = 3
    say("rank-deficient walker through the orthogonalisation")
    W = walkers(s, nw, seed=11, noise=0.3)
    W[1, 1, :, 5] = W[1, 1, :, 2]  # walker 1, spin down: column 5 repeats column 2
    g = m.AfqmcSys(s, nw, "fp64_dense")
    g.set(nat.BUF_W, W)
    g.orthogonalize()
    Wg = g.W
    g.close()
    Wo = orc.orthogonalize(W)
    N = s.naea
    for w in range(nw):
        for sp in range(2):
            Q = Wg[w, sp]
            assert np.abs(Q.T @ Q.conj() - np.eye(N)).max() < 1e-12
    # the full-rank matrices agree entirely, the deficient one up to its
    # dependent column (beyond it the reflectors act on rounding noise)
    full = [(w, sp) for w in range(nw) for sp in range(2) if (w, sp) != (1, 1)]
    for w, sp in full:
        assert rel(Wg[w, sp], Wo[w, sp]) < 1e-10
    assert rel(Wg[1, 1][:, :5], Wo[1, 1][:, :5]) < 1e-10
