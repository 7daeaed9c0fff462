"""include/mpf_math.h: the double-precision tan/acos/tanh shared by the oracle
and the GPU FP64 solver, against the C library (which Python's math module
calls) over the arguments the update uses: acos on [-1, 1], tan on
[1e-6, pi/2 - 1e-6] and at a0, tanh on [1e-6, 1e6] (src/gpu_kernels_3d.py:24-55)."""
import math

import numpy as np

from oracle import oracle as O


def _ulps(a, b):
    if a == b:
        return 0.0
    return abs(a - b) / (np.spacing(abs(b)))


def test_tan_acos_tanh_within_a_few_ulp():
    L = O.lib()
    rng = np.random.default_rng(1)
    y = rng.uniform(1e-6, math.pi / 2 - 1e-6, 20000)
    assert max(_ulps(L.orc_mpf_tan(v), math.tan(v)) for v in y) <= 4
    c = rng.uniform(-1.0, 1.0, 20000)
    assert max(_ulps(L.orc_mpf_acos(v), math.acos(v)) for v in c) <= 2
    x = 10.0 ** rng.uniform(-6, 6, 20000)
    assert max(_ulps(L.orc_mpf_tanh(v), math.tanh(v)) for v in x) <= 8


def test_special_values():
    L = O.lib()
    assert L.orc_mpf_acos(1.0) == 0.0
    assert L.orc_mpf_acos(-1.0) == math.pi
    assert L.orc_mpf_acos(0.0) == math.pi / 2
    assert L.orc_mpf_tanh(1e6) == 1.0 and L.orc_mpf_tanh(0.0) == 0.0
    assert L.orc_mpf_tanh(-2.0) == -L.orc_mpf_tanh(2.0)
    assert _ulps(L.orc_mpf_tan(math.radians(54.7)), math.tan(math.radians(54.7))) <= 1
    # symmetry tan(pi/2 - y) = 1/tan(y) near the clamp of calc_b
    for v in (1e-6, 1e-3, 0.3):
        assert L.orc_mpf_tan(math.pi / 2 - v) * L.orc_mpf_tan(v) == np.float64(1.0) or \
            abs(L.orc_mpf_tan(math.pi / 2 - v) * L.orc_mpf_tan(v) - 1.0) < 1e-9


def test_calc_a_b_mpf_vs_libm():
    s = dict(zip(O.PARAM_KEYS, [0.0] * 13))
    s.update(a0=math.radians(54.7), delta_a=0.36, mu_a=0.6156, p_round=0.05, ksi=0.05,
             theta_c_rad=math.radians(10.0))
    rng = np.random.default_rng(0)
    for c in np.concatenate([rng.uniform(math.cos(math.radians(10.5)), 1.0, 4000), [1.0, 0.999999999]]):
        a = O.calc_b(c, s, use_libm=False)
        b = O.calc_b(c, s, use_libm=True)
        assert abs(a - b) <= 2e-15 * abs(b)
    for c in np.linspace(-1.0, 1.0, 2001):
        assert abs(O.calc_a(c, s, use_libm=False) - O.calc_a(c, s, use_libm=True)) <= 4e-16
