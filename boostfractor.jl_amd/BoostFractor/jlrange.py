"""Julia floating-point ranges `start:step:stop` reproduced element for element.

BoostFractor's coordinate grids are Julia ranges (`X = -0.5:0.01:0.5`,
src/boost_fractor.jl:30) and its k-space grid is the range
`-kmax:dk:kmax` built in `get_kspace_coords` (src/boost_fractor.jl:42-47).
Julia builds such ranges in one of two ways (Base `(:)(::T, ::T, ::T)` for
IEEE floats):
  * when start, step and stop are exactly representable as small rationals
    (continued-fraction convergents with numerator/denominator <= 2^24), the
    elements are (start_n + i*step_n)/den rounded once;
  * otherwise the elements are start + i*step evaluated in double precision
    with the sum taken exactly (TwicePrecision), i.e. fl(start + fl(i*step)).
Both are reproduced here so grids passed to the engine are the same doubles
the reference would hold.
"""
from fractions import Fraction
import math

import numpy as np

_M24 = 16777216  # maxintfloat(Float32), the bound Julia's rat() uses for Float64


def _rat(x):
    """Julia Base.rat(x) for Float64: continued-fraction convergent a/b with
    float(a)/float(b) == x, giving up when max(|a|,|b|) exceeds 2^24."""
    y = x
    a = d = 1
    b = c = 0
    m = _M24
    while abs(y) <= m:
        f = int(y)  # trunc
        y -= f
        a, c = f * a + c, a
        b, d = f * b + d, b
        if max(abs(a), abs(b)) > m:
            return c, d
        if b != 0 and float(a) / float(b) == x:
            break
        if y == 0:
            y = math.inf
        else:
            y = 1.0 / y
    return a, b


def _isbetween(a, x, b):
    return (a <= x <= b) or (b <= x <= a)


def julia_range(start, step, stop):
    start, step, stop = float(start), float(step), float(stop)
    if step == 0:
        raise ValueError("range step cannot be zero")
    step_n, step_d = _rat(step)
    if step_d != 0 and step_n / step_d == step:
        start_n, start_d = _rat(start)
        stop_n, stop_d = _rat(stop)
        if start_d != 0 and stop_d != 0 and start_n / start_d == start and stop_n / stop_d == stop:
            den = start_d * step_d // math.gcd(start_d, step_d)
            m = 2 ** 53
            if den != 0 and abs(start * den) <= m and abs(step * den) <= m and den % start_d == 0 \
                    and den % step_d == 0:
                sn = int(round(start * den))
                st = int(round(step * den))
                num = den * stop_n - stop_d * sn + st * stop_d
                dd = st * stop_d
                sign = -1 if (num < 0) != (dd < 0) else 1
                length = max(0, sign * (abs(num) // abs(dd)))  # Julia div truncates toward zero
                if _isbetween(start, start + (length - 1) * step, stop + step / 2) and \
                        not _isbetween(start, start + length * step, stop):
                    return np.array([float(Fraction(sn + i * st, den)) for i in range(length)])
    lf = (stop - start) / step
    if lf < 0:
        length = 0
    elif lf == 0:
        length = 1
    else:
        length = int(round(lf)) + 1
        stop2 = start + (length - 1) * step
        length -= int(start < stop < stop2) + int(start > stop > stop2)
    return np.array([float(Fraction(start) + Fraction(i * step)) for i in range(length)])
