"""The engine's correctly rounded sin / cos / cube (the values the reference
takes from the C library) against exact evaluations: rational arithmetic for the
cube, a 60-digit Taylor series for sin and cos.  Host code path (CPU)."""
import math
import random
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

from zeroslam_amd.engine import math_probe

getcontext().prec = 60


def exact_sin(x):
    x = Decimal(x)
    t = s = x
    k = 1
    while True:
        t = -t * x * x / ((2 * k) * (2 * k + 1))
        s += t
        k += 1
        if abs(t) < Decimal(10) ** -58:
            return float(s)


def exact_cos(x):
    x = Decimal(x)
    t = s = Decimal(1)
    k = 1
    while True:
        t = -t * x * x / ((2 * k - 1) * (2 * k))
        s += t
        k += 1
        if abs(t) < Decimal(10) ** -58:
            return float(s)


def samples(n, seed):
    rnd = random.Random(seed)
    return np.array([rnd.choice((1, -1)) * 10 ** rnd.uniform(-20, 1.3) for _ in range(n)])


def test_sin_cos_correctly_rounded():
    x = samples(3000, 5)
    s = math_probe(0, x, on_device=False)
    c = math_probe(1, x, on_device=False)
    bad_s = [v for v, r in zip(x, s) if exact_sin(v) != r]
    bad_c = [v for v, r in zip(x, c) if exact_cos(v) != r]
    assert not bad_s and not bad_c, (bad_s[:3], bad_c[:3])


def test_sin_cos_special_values():
    x = np.array([0.0, -0.0, 1e-300, 1e-8, math.pi / 4, math.pi / 2, math.pi, 3 * math.pi / 2, 1e4])
    s = math_probe(0, x, on_device=False)
    c = math_probe(1, x, on_device=False)
    assert np.array_equal(np.signbit(s[:2]), [False, True]) and s[0] == 0.0
    for v, a, b in zip(x[2:-1], s[2:-1], c[2:-1]):
        assert a == exact_sin(v) and b == exact_cos(v)
    # large argument: argument reduction by multiples of pi/2 (Taylor reference unusable)
    assert s[-1] == math.sin(1e4) and c[-1] == math.cos(1e4)


def test_cube_correctly_rounded():
    rnd = random.Random(7)
    y = np.array([rnd.uniform(-1, 1) for _ in range(20000)] + [0.5, -0.25, 1.0, -1.0, 0.0])
    c = math_probe(2, y, on_device=False)
    bad = [v for v, r in zip(y, c) if float(Fraction(float(v)) ** 3) != r]
    assert not bad, bad[:3]


def test_agreement_with_c_library():
    """Where the C library is correctly rounded the two agree; count the rest."""
    x = samples(20000, 9)
    small = np.abs(x) < 0.07  # LM rotation increments live here
    s = math_probe(0, x, on_device=False)
    c = math_probe(1, x, on_device=False)
    libm_s = np.array([math.sin(v) for v in x])
    libm_c = np.array([math.cos(v) for v in x])
    assert np.array_equal(s[small], libm_s[small]) and np.array_equal(c[small], libm_c[small])
    mism = np.count_nonzero(s != libm_s) + np.count_nonzero(c != libm_c)
    assert mism <= 0.002 * len(x)
