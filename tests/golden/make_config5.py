"""Writes tests/golden/config5_oracle.npz: the oracle's reference values for
the uneven multi-system batch (miniafqmc_amd.system.MULTI: M = 96 ... 640,
N = M/4, K = 4M, 15-40 % density, 64-512 walkers).

For every member: walkers 0 and 1 of its run (the counter-based field
generator gives them the draws they have among all of the member's walkers),
four substeps from the trial with the orthogonalisation after substep 2
(northo = 2), the half-rotated Cholesky matrix with every entry kept (what
the dense modes contract), then the local energy in the Cholesky form.  The
fixture keeps per walker: E_L, the overlaps, the hybrid energies, and 512
entries of the walker matrices at seeded positions with their indices.
Walkers couple only through Eshift, which scales the weights alone, so these
values do not depend on the other walkers.

Run from the repository root after building the oracle and the library:

    python tests/golden/make_config5.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import miniafqmc_amd as m  # noqa: E402
import oracle_lib as orc  # noqa: E402
from miniafqmc_amd.system import MULTI  # noqa: E402

NW, NSUB, NORTHO, NSAMP = 2, 4, 2, 512


def main():
    orc.set_threads(0)
    out = {"nsub": NSUB, "northo": NORTHO, "nw_sample": NW}
    for i in range(len(MULTI)):
        t0 = time.time()
        s, nw, seed = m.multi_system(i)
        M, N, K = s.nmo, s.naea, s.nchol
        spT = orc.halfrotate(s, cutoff=-1.0)
        Lt = np.ascontiguousarray(spT.dense() / np.sqrt(s.dt))
        x = m.fields_block(seed, K, NW, 0, NSUB, w0=0, wtot=nw)
        eo, W, Wd, es, tm, info = orc.run(s, NW, 1, NSUB, NORTHO, x, energy_mode=1, spT=spT, Lt=Lt)
        assert info == 0, info
        del spT, Lt
        r = np.random.default_rng(7000 + i)
        idx = r.choice(2 * M * N, NSAMP, replace=False).astype(np.int64)
        flat = W.reshape(NW, 2 * M * N)
        out[f"s{i}_shape"] = np.array([M, N, K, nw], np.int64)
        out[f"s{i}_E_L"] = Wd[:, 0].copy()
        out[f"s{i}_ovlp"] = Wd[:, 2:4].copy()
        out[f"s{i}_hyb"] = Wd[:, 4:6].copy()
        out[f"s{i}_widx"] = idx
        out[f"s{i}_wval"] = flat[:, idx].copy()
        out[f"s{i}_wmax"] = np.abs(flat).max(axis=1)
        print(f"system {i}: M={M} N={N} K={K} walkers={nw}: E_L {Wd[:, 0].real} ({time.time() - t0:.1f}s)",
              flush=True)
        del s, W, Wd
    np.savez_compressed(os.path.join(HERE, "config5_oracle.npz"), **out)


if __name__ == "__main__":
    main()
