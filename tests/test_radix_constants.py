"""The FP32 odd-radix DFT constants compiled into the kernels (kOddRoots in csrc/qpd_device.hpp)
equal the FP32 values of the host's twiddle table, for every odd radix the kernels use.

The kernels evaluate cos and sin of 2 pi m / R at compile time with a Taylor series in double
(argument folded to |x| <= pi/2) and round to float; the host table (qpd.hip unit_root) takes
cosl / sinl of 2 pi m / R in long double, rounds to double and then to float.  This restates
both in Python (IEEE double arithmetic in the same order as the constant evaluation; numpy
longdouble for the host) and requires equal float32 values.
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
