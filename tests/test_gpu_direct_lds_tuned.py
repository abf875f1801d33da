"""Direct LDS staging (a ring of LDS stages for both split operands) with the
tile shapes the autotuner picks: the tuned split-BF16 walk gives the same
walkers as the tuned register-staged walk (all tile shapes accumulate k in
the same order)."""
import sys

import numpy as np
import pytest

import miniafqmc_amd as m

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("shape", ["N64_M256", "N128_M512"])
def test_direct_lds_tuned_tiles_match_register_path(require_gpu, shape):
    s, p = m.preset_system(shape)
    nw = 128
    x = m.fields_block(p["seed"], s.nchol, nw, 0, 3)
    out = []
    for direct in (0, 2):
        say(f"tuned direct LDS {shape} level={direct}")
        g = m.AfqmcSys(s, nw, "mixed_bf16", dense_exchange=False)
        g.set_direct_lds(direct)
        g.stage_fields(x)
        g.tune(s.dt)
        g.init_walkers()
        g.calculate_mixed_density_matrix()
        g.run(1, 3, 2, s.dt, energy=False)
        out.append(g.W)
        g.close()
    d = np.abs(out[0] - out[1]).max() / np.abs(out[0]).max()
    say(f"{shape}: max rel diff {d:.3e}")
    assert d == 0.0
