"""The per-pair flop count that prices the field-sum roofline.

One gausslet x point term of sum_gaussian_modes (cfields.pyx:102-114 and
calc_mode_U, cfields.pyx:118-150) is evaluated here on counting numbers:
every real +, -, x and / is one flop, complex operations are expanded into
their real ones (complex x complex 6, complex x real 2, complex +- complex
2, complex +- real 1, complex / complex 11 by the textbook formula,
complex / real 2, multiplication by i 0), and exp, sin/cos and sqrt are
counted apart.  What depends only on the gausslet is hoisted out of the
pair and not counted: the unit vectors E, H, the direction, the origin, the
scaled A, B, C and A + C, detG0, k·n, the phase, and
inv_root_area·sqrt(i).  The total is the integer bench.py prices with."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class Tally(object):
    def __init__(self):
        self.flops = 0
        self.special = {}

    def op(self, n=1):
        self.flops += n

    def fn(self, name):
        self.special[name] = self.special.get(name, 0) + 1


T = Tally()


class R(object):
    """A real value whose arithmetic is counted (the values themselves are
    not needed)."""

    def _bin(self, other):
        if isinstance(other, C):
            return NotImplemented
        T.op()
        return R()

    __add__ = __radd__ = __sub__ = __rsub__ = __mul__ = __rmul__ = __truediv__ = __rtruediv__ = _bin


def as_c(v):
    return v if isinstance(v, C) else C(v, None)


class C(object):
    """A complex value; im None marks a real operand (so complex x real is
    two products, not six)."""

    def __init__(self, re, im):
        self.re, self.im = re, im

    def __add__(self, o):
        o = as_c(o)
        if o.im is None or self.im is None:
            T.op(1)
        else:
            T.op(2)
        return C(R(), R())

    __radd__ = __add__
    __sub__ = __add__
    __rsub__ = __add__

    def __mul__(self, o):
        o = as_c(o)
        T.op(2 if (o.im is None or self.im is None) else 6)
        return C(R(), R())

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = as_c(o)
        T.op(2 if o.im is None else 11)
        return C(R(), R())

    def times_i(self):
        return C(R(), R())


def cexp(w):
    T.fn("exp")
    T.fn("sincos")
    T.op(2)   # exp(re)·cos(im), exp(re)·sin(im)
    return C(R(), R())


def csqrt(w):
    # principal square root: h = hypot(re, im) (2 products, 1 sum, 1 sqrt);
    # t = sqrt((|re| + h)/2) (1 sum, 1 product, 1 sqrt); im/(2t) (1 product,
    # 1 quotient)
    T.fn("sqrt")
    T.fn("sqrt")
    T.op(6)
    return C(R(), R())


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def one_pair():
    """The reference's per-pair expressions, operand for operand."""
    # hoisted per gausslet
    origin = [R(), R(), R()]
    E = [R(), R(), R()]
    H = [R(), R(), R()]
    d = [R(), R(), R()]
    A, B, Cc, AplusC, detG0, kz = (C(R(), R()) for _ in range(6))
    phase = R()
    norm = C(R(), R())              # inv_root_area * rootI
    E1_amp, E2_amp = C(R(), R()), C(R(), R())
    out = [C(R(), R()) for _ in range(3)]
    # per pair (cfields.pyx:102-114)
    point = [R(), R(), R()]
    pt = [point[i] - origin[i] for i in range(3)]
    # calc_mode_U (cfields.pyx:133-150)
    x, y, z = dot(pt, E), dot(pt, H), dot(pt, d)
    denom = (1 + (AplusC * z) + (z * z) * detG0) * 2
    zdet = detG0 * z
    AA = (A + zdet) / denom
    CC = (Cc + zdet) / denom
    arg = AA * (x * x) + z
    arg = arg + (B * (2 * (x * y))) / denom
    arg = arg + CC * (y * y)
    U = cexp((phase + kz * arg).times_i())
    zA, zC, zB = A * z, Cc * z, B * z
    U = U / csqrt((((zA + 1) * (zC + 1)) - zB * zB).times_i())
    U = U * norm
    E1 = E1_amp * U
    E2 = E2_amp * U
    for c in range(3):
        out[c] = out[c] + (E1 * E[c] + E2 * H[c])


def test_field_pair_flops_pinned():
    T.flops, T.special = 0, {}
    one_pair()
    assert T.special == {"exp": 1, "sincos": 1, "sqrt": 2}
    with open(os.path.join(ROOT, "bench.py")) as f:
        pinned = int(re.search(r"^FIELD_FLOPS_PER_PAIR = (\d+)", f.read(), re.M).group(1))
    assert T.flops == pinned == 172


def test_counting_rules():
    """The counter's own conventions on operations of known cost."""
    T.flops = 0
    C(R(), R()) * C(R(), R())
    assert T.flops == 6
    T.flops = 0
    C(R(), R()) * R()
    C(R(), R()) + 1
    assert T.flops == 3
    T.flops = 0
    C(R(), R()) / C(R(), R())
    C(R(), R()) / 2
    assert T.flops == 13
    T.flops = 0
    dot([R(), R(), R()], [R(), R(), R()])
    assert T.flops == 5
