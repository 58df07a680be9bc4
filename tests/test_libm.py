"""The deterministic exp/log/pow (watergap2_amd/csrc/wg_libm.h) used by the kernels
and the oracle: agreement with the host C library, correct rounding where the two
disagree (checked with 60-digit decimal arithmetic), special cases, and -- on the
GPU -- identical bits on the device."""
import ctypes as C
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest


def samples(n=20000, seed=3):
    rng = np.random.default_rng(seed)
    x_exp = rng.uniform(-700, 700, n)
    x_exp[: n // 4] = rng.uniform(-20, 20, n // 4)
    base = rng.uniform(0, 3, n)
    base[: n // 5] = rng.uniform(200, 350, n // 5)  # pow(temp_K, 4)
    y = rng.uniform(0, 8, n)
    y[::4] = 0.5
    y[1::4] = 4.0
    y[2::7] = 2.0 / 3.0
    x_log = np.exp(rng.uniform(-50, 50, n))
    return x_exp, base, y, x_log


def ulps(a, b):
    ia = np.array([a]).view(np.int64)[0]
    ib = np.array([b]).view(np.int64)[0]
    return abs(int(ia) - int(ib))


def exact_exp(x):
    getcontext().prec = 60
    return Decimal(float(x)).exp()


def exact_pow(x, y):
    getcontext().prec = 60
    return (Decimal(float(y)) * Decimal(float(x)).ln()).exp()


def err_ulp(v, exact):
    return float(abs(Decimal(float(v)) - exact) / Decimal(math.ulp(float(exact))))


def test_against_glibc_and_correct_rounding(oracle):
    lib = oracle.load("wg")
    x_exp, base, y, x_log = samples()
    diffs = {"exp": 0, "pow": 0, "log": 0}
    for i in range(len(x_exp)):
        a, b = lib.wgo_exp(x_exp[i]), math.exp(x_exp[i])
        if a != b:
            diffs["exp"] += 1
            assert ulps(a, b) == 1
            assert err_ulp(a, exact_exp(x_exp[i])) <= 0.5  # ours is the correctly rounded one
        a, b = lib.wgo_pow(base[i], y[i]), math.pow(base[i], y[i])
        if a != b:
            diffs["pow"] += 1
            assert ulps(a, b) == 1
            assert err_ulp(a, exact_pow(base[i], y[i])) <= 0.5
        a, b = lib.wgo_log(x_log[i]), math.log(x_log[i])
        if a != b:
            diffs["log"] += 1
            assert ulps(a, b) <= 1
    n = len(x_exp)
    assert all(v < 0.005 * n for v in diffs.values()), diffs


def test_special_cases(oracle):
    lib = oracle.load("wg")
    assert lib.wgo_pow(1.0, 3.7) == 1.0       # S = Smax gives R = Peff exactly (daily.cpp:1126)
    assert lib.wgo_pow(0.0, 2.5) == 0.0       # empty store: no outflow
    assert lib.wgo_pow(0.25, 0.5) == 0.5
    assert lib.wgo_pow(2.0, 10.0) == 1024.0
    assert lib.wgo_pow(3.0, 4.0) == 81.0
    assert lib.wgo_exp(0.0) == 1.0
    assert lib.wgo_exp(-800.0) == 0.0
    assert math.isinf(lib.wgo_exp(710.0))
    assert lib.wgo_exp(1.0) == math.e
    assert lib.wgo_log(1.0) == 0.0
    assert lib.wgo_log(math.e) == 1.0


@pytest.mark.gpu
def test_device_bits_equal_host(wg, oracle):
    from watergap2_amd import _abi
    lib = _abi.load_library()
    host = oracle.load("wg")
    x_exp, base, y, x_log = samples(100000, seed=9)
    for fn, (xs, ys, ref) in {
        0: (x_exp, None, lambda a, b: host.wgo_exp(a)),
        1: (base, y, lambda a, b: host.wgo_pow(a, b)),
        2: (base * 1e3, None, lambda a, b: math.sqrt(a)),
        3: (x_log, None, lambda a, b: host.wgo_log(a)),
        4: (np.random.default_rng(4).random(20000), None, lambda a, b: oracle.sincos2pi(a)[0]),
        5: (np.random.default_rng(5).random(20000), None, lambda a, b: oracle.sincos2pi(a)[1]),
    }.items():
        xs = np.ascontiguousarray(xs)
        ys_ = np.ascontiguousarray(ys) if ys is not None else None
        out = np.zeros_like(xs)
        rc = lib.wg_libm_eval(fn, len(xs), _abi.ptr(xs, _abi.P_d),
                              _abi.ptr(ys_, _abi.P_d) if ys_ is not None else None,
                              _abi.ptr(out, _abi.P_d))
        assert rc == 0
        expect = np.array([ref(xs[i], ys_[i] if ys_ is not None else 0.0) for i in range(len(xs))])
        mism = np.flatnonzero(out.view(np.int64) != expect.view(np.int64))
        assert len(mism) == 0, (fn, len(mism), xs[mism[:3]], out[mism[:3]], expect[mism[:3]])


def test_sincos2pi_accuracy(oracle):
    """sin and cos of 2 pi u within 3e-16 of an extended-precision evaluation
    (x87 long double: the angle 2 pi u carries no double rounding there) on random
    and on quadrant-boundary arguments."""
    rng = np.random.default_rng(6)
    us = np.concatenate([rng.random(20000), np.arange(0, 1, 1 / 64), [0.5 - 2 ** -40, 0.25 + 2 ** -45]])
    got = np.array([oracle.sincos2pi(u) for u in us])
    two_pi = np.longdouble(2) * np.longdouble("3.14159265358979323846264338327950288")
    ang = two_pi * us.astype(np.longdouble)
    assert np.finfo(np.longdouble).eps < 1e-18
    np.testing.assert_allclose(got[:, 0], np.sin(ang).astype(np.float64), rtol=0, atol=3e-16)
    np.testing.assert_allclose(got[:, 1], np.cos(ang).astype(np.float64), rtol=0, atol=3e-16)
