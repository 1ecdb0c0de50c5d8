"""Deterministic log/exp/sin/cos (abeille_amd/csrc/detmath.h) against glibc:
at most one ulp over the ranges the transport path uses, exact special
values, and identical results from the oracle's and the device header's
builds (same source, -ffp-contract=off)."""
import ctypes as C
import math

import numpy as np

from oracle import pyoracle
import pyhd

O = pyoracle.lib()
H = pyhd.lib()
H.hd_log.argtypes = [C.c_double]
H.hd_log.restype = C.c_double
H.hd_exp.argtypes = [C.c_double]
H.hd_exp.restype = C.c_double
H.hd_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]


def ulp_diff(a, b):
    return abs(np.array([a]).view(np.int64)[0] - np.array([b]).view(np.int64)[0])


def test_log_one_ulp():
    rng = np.random.default_rng(1)
    xs = np.concatenate([1.0 - rng.random(20000), rng.random(2000) * 1e-300,
                         np.exp(rng.uniform(-700, 700, 2000)), [2.0 ** -53, 0.5, 2.0, 1e-310]])
    worst = 0
    for x in xs:
        a = O.orc_log(1, float(x))
        assert a == H.hd_log(float(x))
        worst = max(worst, ulp_diff(a, math.log(x)))
    assert worst <= 1


def test_log_special_values():
    assert O.orc_log(1, 1.0) == 0.0
    assert O.orc_log(1, 0.0) == -math.inf
    assert math.isnan(O.orc_log(1, -1.0))
    assert O.orc_log(1, math.inf) == math.inf
    assert abs(O.orc_log(1, math.e) - 1.0) <= 2.3e-16


def test_sincos_one_ulp():
    rng = np.random.default_rng(2)
    xs = np.concatenate([2 * math.pi * rng.random(20000), [0.0, math.pi / 4, math.pi / 2,
                                                           math.pi, 2 * math.pi]])
    worst = 0
    s, c = C.c_double(), C.c_double()
    for x in xs:
        a, b = O.orc_sin(1, float(x)), O.orc_cos(1, float(x))
        H.hd_sincos(float(x), C.byref(s), C.byref(c))
        assert (a, b) == (s.value, c.value)
        if abs(math.sin(x)) > 1e-3:
            worst = max(worst, ulp_diff(a, math.sin(x)))
        if abs(math.cos(x)) > 1e-3:
            worst = max(worst, ulp_diff(b, math.cos(x)))
        assert abs(a * a + b * b - 1.0) < 4e-16
    assert worst <= 1


def test_exp_one_ulp():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-745, 709.7, 20000), rng.uniform(-50, 5, 20000),
                         rng.uniform(-1e-3, 1e-3, 2000), [0.0, -0.0, 1.0, -1.0, 0.5 * math.log(2),
                                                          709.78, -708.3, -720.0, -744.5]])
    worst = 0
    for x in xs:
        a = O.orc_exp(1, float(x))
        assert a == H.hd_exp(float(x))
        b = math.exp(x)
        if b < 2.2250738585072014e-308:        # subnormal results: two rounding steps
            assert abs(a - b) <= 2 * 5e-324 * 1.0000001, (x, a, b)
            continue
        worst = max(worst, ulp_diff(a, b))
    assert worst <= 1


def test_exp_special_values():
    assert O.orc_exp(1, 0.0) == 1.0
    assert O.orc_exp(1, 710.0) == math.inf
    assert O.orc_exp(1, -746.0) == 0.0
    assert math.isnan(O.orc_exp(1, math.nan))
    assert abs(O.orc_exp(1, 1.0) - math.e) <= 4.5e-16
