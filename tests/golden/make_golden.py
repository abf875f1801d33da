"""Writes tests/golden/walker_step_m16.npz: one small walker-step case, inputs
and the oracle's outputs phase by phase plus a 6-substep trajectory.

System: 16 orbitals, 4+4 electrons, 32 Cholesky vectors at 50% density with
Vakbl, 4 walkers perturbed away from the trial, dt = 0.005.  Run from the
repository root after building the oracle and the library:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import miniafqmc_amd as m  # noqa: E402
import oracle_lib as orc  # noqa: E402

SEED = 24061
NW = 4


def main():
    s = m.make_system(16, 4, 32, density=0.5, sigma_L=0.05, sigma_h=1.0, dt=0.005, seed=SEED, with_vakbl=True)
    r = np.random.default_rng(SEED)
    W = np.zeros((NW, 2, s.nmo, s.naea), complex)
    W[:, 0] = s.trial_alpha.conj()
    W[:, 1] = s.trial_beta.conj()
    W += 0.1 * (r.standard_normal(W.shape) + 1j * r.standard_normal(W.shape))
    Wd = np.zeros((NW, 8), complex)
    Wd[:, 1] = 1.0
    Wd0 = Wd.copy()
    x = m.fields(SEED, s.nchol, NW, 0)

    spT = orc.halfrotate(s, cutoff=1e-6)
    G, _ = orc.calculate_mixed_density_matrix(s, W, Wd)
    vb = orc.get_vbias(spT, G)
    X, hw = orc.fields(vb, x)
    vhs = orc.get_vhs(s.spvn, X)
    Wp = orc.propagate(s, W, vhs)
    Wd2 = Wd.copy()
    Wd2[:, 5] = Wd2[:, 4]
    Wd2[:, 6:8] = Wd2[:, 2:4]
    orc.calculate_overlaps(s, Wp, Wd2)
    es = orc.weights(s.dt, hw, Wd2, 0.0)
    Wo = orc.orthogonalize(Wp)
    Wd3 = Wd2.copy()
    G3, _ = orc.calculate_mixed_density_matrix(s, Wo, Wd3)
    eav = orc.calculate_energy(s, Wd3, G3)

    nsub, northo = 6, 2
    xr = m.fields_block(SEED, s.nchol, NW, 0, nsub)
    eo, Wr, Wdr, esr, _, _ = orc.run(s, NW, 1, nsub, northo, xr, energy_mode=0, spT=spT)

    out = dict(
        nmo=s.nmo, naea=s.naea, nchol=s.nchol, dt=s.dt, seed=SEED,
        trial_alpha=s.trial_alpha, trial_beta=s.trial_beta, haj=s.haj, propg1=s.propg1, h1=s.h1,
        spvn_rowptr=s.spvn.rowptr, spvn_col=s.spvn.col, spvn_val=s.spvn.val,
        vakbl_rowptr=s.vakbl.rowptr, vakbl_col=s.vakbl.col, vakbl_val=s.vakbl.val,
        spvnT_rowptr=spT.rowptr, spvnT_col=spT.col, spvnT_val=spT.val,
        W=W, Wd=Wd0, x=x,
        G=G, ovlp=Wd[:, 2:4], vbias=vb, X=X, hybridW=hw, vHS=vhs, W_prop=Wp, Wd_weights=Wd2, eshift=es,
        W_ortho=Wo, G_ortho=G3, E_L=Wd3[:, 0], Eav=eav,
        run_nsub=nsub, run_northo=northo, run_x=xr, run_eav=eo, run_W=Wr, run_Wd=Wdr, run_eshift=esr,
    )
    path = os.path.join(HERE, "walker_step_m16.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
