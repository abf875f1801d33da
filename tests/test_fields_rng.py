"""The auxiliary-field generator (afqmc_synth_fields, host side of the
counter-based generator the fields kernel also runs on the GPU).

The reference draws x ~ N(0,1) by Box-Muller from a Mersenne-Twister stream
(src/Utilities/RandomGenerator.h:50-62, called at
src/miniapps/miniafqmc_base.cpp:282); this build's stream is counter-based,
so its pins are mathematical: the uniform pair of each draw is recomputed
here from the published splitmix64 finaliser, and the draw must equal
sqrt(-2 ln u1) (cos, sin)(2 pi u2) evaluated with numpy's libm to 1e-14
(a wrong polynomial coefficient, quadrant or index mapping fails this), and
the sample moments must be those of N(0,1).
"""
import numpy as np

import miniafqmc_amd as m

MASK = (1 << 64) - 1


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def u01(seed, stream, idx):
    h = mix64(mix64((seed * 0x2545F4914F6CDD1D + stream) & MASK) ^ ((idx * 0xD1342543DE82EF95) & MASK))
    return ((h >> 11) + 1) * (1.0 / 9007199254740992.0)


def stream(substep, w):
    return ((5 << 56) ^ (substep << 24) ^ w) & MASK


def expected(seed, substep, w, n):
    j = n // 2
    u1, u2 = u01(seed, stream(substep, w), 2 * j), u01(seed, stream(substep, w), 2 * j + 1)
    r = np.sqrt(-2.0 * np.log(u1))
    return r * (np.cos(2 * np.pi * u2) if n % 2 == 0 else np.sin(2 * np.pi * u2))


def test_draws_equal_box_muller_of_the_counter_uniforms():
    seed, K, nw = 20261234, 37, 5
    for sub in (0, 3, 1000):
        x = m.fields(seed, K, nw, sub)
        for w in range(nw):
            for n in range(K):
                e = expected(seed, sub, w, n)
                assert abs(x[n, w] - e) <= 1e-14 * max(1.0, abs(e)), (sub, w, n, x[n, w], e)


def test_walker_subset_reproduces_columns():
    seed, K = 99, 64
    full = m.fields(seed, K, 40, 7)
    part = m.fields(seed, K, 9, 7, w0=13, wtot=40)
    assert np.array_equal(full[:, 13:22], part)


def test_moments_and_tails():
    x = m.fields_block(5, 2048, 256, 0, 4).ravel()  # 2M draws
    n = x.size
    assert abs(x.mean()) < 5 / np.sqrt(n)
    assert abs(x.var() - 1.0) < 5 * np.sqrt(2.0 / n)
    assert abs((x ** 3).mean()) < 5 * np.sqrt(15.0 / n)
    assert abs((x ** 4).mean() - 3.0) < 5 * np.sqrt(96.0 / n)
    # tail fraction beyond 3 sigma: 2 (1 - Phi(3)) = 0.0026998
    p = 0.0026997960632601866
    assert abs(np.mean(np.abs(x) > 3.0) - p) < 5 * np.sqrt(p * (1 - p) / n)
