"""The odd-radix DFT constants compiled into the kernels (kOddRoots and kOddRoots64 in
csrc/qpd_device.hpp) equal the FP32 and FP64 values of the host's twiddle table, for every odd
radix the kernels use.

The kernels evaluate cos and sin of 2 pi m / R at compile time with a Taylor series in double
(argument folded to |x| <= pi/2) and round to float; the host table (qpd.hip unit_root) takes
cosl / sinl of 2 pi m / R in long double, rounds to double and then to float.  This restates
both in Python (IEEE double arithmetic in the same order as the constant evaluation; numpy
longdouble for the host) and requires equal float32 values.  The FP64 constants are evaluated
in double-double arithmetic and rounded to double; they are restated the same way and must equal
the host's double values.
"""

import numpy as np
import pytest

RADICES = (3, 5, 7, 9, 11, 13, 15, 17)


def taylor_cos(x):
    term = sum_ = 1.0
    for k in range(1, 14):
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k))
        sum_ += term
    return sum_


def taylor_sin(x):
    term = sum_ = x
    for k in range(1, 14):
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0))
        sum_ += term
    return sum_


def compiled(R, m):
    pi = 3.14159265358979323846264338327950288
    past = 4 * m > R
    x = pi * float(R - 2 * m) / float(R) if past else 2.0 * pi * float(m) / float(R)
    c = -taylor_cos(x) if past else taylor_cos(x)
    return np.float32(c), np.float32(taylor_sin(x))


def host_table(R, m):
    if np.finfo(np.longdouble).nmant < 60:
        pytest.skip("numpy longdouble is not x87 extended precision here")
    two_pi = np.longdouble("6.283185307179586476925286766559005768")
    t = two_pi * np.longdouble(m % R) / np.longdouble(R)
    return np.float32(np.float64(np.cos(t))), np.float32(np.float64(np.sin(t)))


@pytest.mark.parametrize("R", RADICES)
def test_odd_radix_constants_match_host_table(R):
    for m in range(R // 2 + 1):
        c, s = compiled(R, m)
        hc, hs = host_table(R, m)
        assert c == hc and s == hs, (R, m, c, hc, s, hs)
        # and both are cos / sin of 2 pi m / R to float rounding
        assert abs(float(c) - np.cos(2 * np.pi * m / R)) <= 6e-8
        assert abs(float(s) - np.sin(2 * np.pi * m / R)) <= 6e-8


# --- the FP64 constants (kOddRoots64): double-double evaluation, rounded to double -----------

def _two_sum(a, b):
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _fast_two_sum(a, b):
    s = a + b
    return s, b - (s - a)


def _split(a):
    t = 134217729.0 * a
    hi = t - (t - a)
    return hi, a - hi


def _two_prod(a, b):
    p = a * b
    ah, al = _split(a)
    bh, bl = _split(b)
    return p, ((ah * bh - p) + ah * bl + al * bh) + al * bl


def _dd_add(a, b):
    s = _two_sum(a[0], b[0])
    t = _two_sum(a[1], b[1])
    s = _fast_two_sum(s[0], s[1] + t[0])
    return _fast_two_sum(s[0], s[1] + t[1])


def _dd_mul(a, b):
    p = _two_prod(a[0], b[0])
    return _fast_two_sum(p[0], p[1] + (a[0] * b[1] + a[1] * b[0]))


def _dd_div(a, b):
    q1 = a[0] / b
    p = _two_prod(q1, b)
    r = _dd_add(a, (-p[0], -p[1]))
    return _fast_two_sum(q1, r[0] / b)


def _dd_taylor(x, odd):
    x2 = _dd_mul(x, x)
    term = x if odd else (1.0, 0.0)
    total = term
    for k in range(1, 28):
        d = (2.0 * k) * (2.0 * k + 1.0) if odd else (2.0 * k - 1.0) * (2.0 * k)
        term = _dd_div(_dd_mul(term, (-x2[0], -x2[1])), d)
        total = _dd_add(total, term)
    return total[0] + total[1]


def compiled64(R, m):
    pi = (3.141592653589793116, 1.2246467991473532e-16)
    past = 4 * m > R
    x = _dd_div(_dd_mul(pi, (float(R - 2 * m if past else 2 * m), 0.0)), float(R))
    c = _dd_taylor(x, False)
    return (-c if past else c), _dd_taylor(x, True)


def host_table64(R, m):
    if np.finfo(np.longdouble).nmant < 60:
        pytest.skip("numpy longdouble is not x87 extended precision here")
    two_pi = np.longdouble("6.283185307179586476925286766559005768")
    t = two_pi * np.longdouble(m % R) / np.longdouble(R)
    return float(np.float64(np.cos(t))), float(np.float64(np.sin(t)))


@pytest.mark.parametrize("R", RADICES)
def test_odd_radix_constants_fp64_match_host_table(R):
    """kOddRoots64 (csrc/qpd_device.hpp) equals the host's FP64 twiddle table exactly."""
    for m in range(R // 2 + 1):
        c, s = compiled64(R, m)
        hc, hs = host_table64(R, m)
        assert c == hc and s == hs, (R, m, c, hc, s, hs)
        # numpy's double cos / sin of the rounded double argument, to a few ulp
        assert abs(c - np.cos(2 * np.pi * m / R)) <= 1e-15
        assert abs(s - np.sin(2 * np.pi * m / R)) <= 1e-15


def test_planted_fp64_constant_error_is_caught():
    """A double Taylor series (no double-double) misses the host table for some constant."""
    misses = 0
    for R in RADICES:
        for m in range(R // 2 + 1):
            past = 4 * m > R
            pi = 3.14159265358979323846264338327950288
            x = pi * float(R - 2 * m) / float(R) if past else 2.0 * pi * float(m) / float(R)
            c = -taylor_cos(x) if past else taylor_cos(x)
            hc, hs = host_table64(R, m)
            misses += (c != hc) + (taylor_sin(x) != hs)
    assert misses > 0
