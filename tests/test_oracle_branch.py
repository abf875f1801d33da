"""Comb population control in the oracle (orc_branch).  The reference only
marks where branching would go (src/miniapps/miniafqmc_base.cpp:346), so the
oracle is pinned by a comb worked out by hand and by the comb's invariants:
walker i is copied floor or ceil of nw*w_i/W times, the population stays nw,
parents come in index order, zero-weight walkers are dropped, the total
weight is kept, and over u uniform the mean number of copies is nw*w_i/W.
"""
import numpy as np
import pytest

import oracle_lib as orc


def walkers(nw, M=3, N=2, seed=0):
    r = np.random.default_rng(seed)
    W = r.standard_normal((nw, 2, M, N)) + 1j * r.standard_normal((nw, 2, M, N))
    Wd = r.standard_normal((nw, 8)) + 1j * r.standard_normal((nw, 8))
    return W, Wd


def test_hand_worked_comb():
    W, Wd = walkers(3)
    Wd[:, 1] = [1.0, 0.0, 3.0]
    W0, Wd0 = W.copy(), Wd.copy()
    # W = 4, teeth at (0.5 + j) * 4/3 = 0.667, 2.0, 3.333; cumulative weights 1, 1, 4
    st, parent = orc.branch(W, Wd, 0.5)
    assert st == 0
    np.testing.assert_array_equal(parent, [0, 2, 2])
    np.testing.assert_array_equal(W, W0[[0, 2, 2]])
    np.testing.assert_array_equal(np.delete(Wd, 1, axis=1), np.delete(Wd0[[0, 2, 2]], 1, axis=1))
    np.testing.assert_allclose(Wd[:, 1], 4.0 / 3.0, rtol=0, atol=1e-15)
    # a tooth on a boundary goes to the next walker (cum > t): u = 0.75 -> t = 1, 2.333, 3.667
    W, Wd = walkers(3)
    Wd[:, 1] = [1.0, 0.0, 3.0]
    st, parent = orc.branch(W, Wd, 0.75)
    np.testing.assert_array_equal(parent, [2, 2, 2])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_comb_invariants(seed):
    nw = 997
    W, Wd = walkers(nw, seed=seed)
    r = np.random.default_rng(seed + 100)
    w = r.exponential(1.0, nw)
    w[r.random(nw) < 0.2] = 0.0      # phaseless-killed walkers
    w[5] = -0.3                      # a negative weight counts as zero
    Wd[:, 1] = w
    wpos = np.maximum(w, 0.0)
    Wt = wpos.sum()
    st, parent = orc.branch(W, Wd, r.random())
    assert st == 0
    assert np.all(np.diff(parent) >= 0)
    counts = np.bincount(parent, minlength=nw)
    x = nw * wpos / Wt
    assert counts.sum() == nw
    assert np.all((counts == np.floor(x)) | (counts == np.ceil(x)))
    assert np.all(counts[wpos == 0] == 0)
    np.testing.assert_allclose(Wd[:, 1].real.sum(), Wt, rtol=1e-13)


def test_mean_copies_over_u():
    nw = 64
    r = np.random.default_rng(9)
    w = r.exponential(1.0, nw)
    us = (np.arange(2000) + 0.5) / 2000
    mean = np.zeros(nw)
    for u in us:
        W, Wd = walkers(nw, M=1, N=1)
        Wd[:, 1] = w
        st, parent = orc.branch(W, Wd, u)
        mean += np.bincount(parent, minlength=nw)
    mean /= len(us)
    np.testing.assert_allclose(mean, nw * w / w.sum(), atol=2e-3)


@pytest.mark.parametrize("bad", ["zero", "nan"])
def test_rejects_degenerate_weights(bad):
    W, Wd = walkers(5)
    Wd[:, 1] = 0.0
    if bad == "nan":
        Wd[2, 1] = np.nan
        Wd[0, 1] = 1.0
    W0, Wd0 = W.copy(), Wd.copy()
    st, _ = orc.branch(W, Wd, 0.3)
    assert st == 1
    np.testing.assert_array_equal(W, W0)
    np.testing.assert_array_equal(Wd[:, [0, 2, 3]], Wd0[:, [0, 2, 3]])
