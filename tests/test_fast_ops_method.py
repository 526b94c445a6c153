"""The short root and quotient sequences of the vertex pass (FastOps in
curvedspacevertexmodel_amd/csrc/svm_geom.h, sqrt_in_range in svm_fp64.h),
restated step for step in exact arithmetic on the CPU: every fma is the
exact a * b + c rounded once (Fraction, then float(), which rounds to
nearest even), every product an IEEE double operation.  The sequences are the
compiler's own division and root refinements without their range wrappers,
so on the device they give the operators' bits wherever the wrappers change
nothing (which the GPU self-check confirms, together with the nearest-double
property of every accepted result).  These tests pin the refinements'
arithmetic:
  * the quotient's final correction (q0 = a y, r = fma(-b, q0, a),
    q = fma(r, y, q0)) gives the correctly rounded a / b from the correctly
    rounded reciprocal y, also for the hard cases (all-ones divisors over
    powers of two, quotients just above a rounding midpoint);
  * two Newton steps keep the correctly rounded reciprocal, and reach it from
    seeds within two units in the last place, except for all-ones divisors,
    whose reciprocal lies 2^-106 above a midpoint: there a seed one unit off
    stops one unit low, so the quotient sequence relies on the seed
    instruction being exact there, as the division operator built from the
    same instructions does;
  * the root refinement gives the correctly rounded root from seeds within
    two units in the last place;
  * the same samples reject a seed of 2^-22 relative accuracy, a refinement
    one step short and a root without its corrections, which shows that they
    can tell a sufficient refinement from an insufficient one."""
import math
import random
from fractions import Fraction

import pytest

def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def seed(x, rng, coarse=False):
    """A seed for x: within two ulps of it, or (coarse) within a relative 2^-22."""
    if coarse:
        return float(Fraction(x) * (1 + Fraction(rng.uniform(-2.0 ** -22, 2.0 ** -22))))
    for _ in range(rng.randrange(3)):
        x = math.nextafter(x, math.inf if rng.random() < 0.5 else -math.inf)
    return x


def rcp(b, y0, steps=2):
    y = y0
    for _ in range(steps):
        e = fma(-b, y, 1.0)
        y = fma(y, e, y)
    return y


def div(a, b, y):
    q0 = a * y
    r = fma(-b, q0, a)
    return fma(r, y, q0)


def sqrt_in_range(x, y, corrections=2):
    g = x * y
    h = y * 0.5
    r = fma(-h, g, 0.5)
    g = fma(g, r, g)
    h = fma(h, r, h)
    for _ in range(corrections):
        d = fma(-g, g, x)
        g = fma(d, h, g)
    return g


def mantissa(rng):
    """A double in [1, 2): random, or with all-ones / all-zero / one-bit fractions."""
    k = rng.randrange(8)
    if k == 0:
        return 2.0 - 2.0 ** -52
    if k == 1:
        return 1.0
    if k == 2:
        return 1.0 + 2.0 ** -52
    return 1.0 + rng.getrandbits(52) * 2.0 ** -52


def quotient_samples(n, rng):
    """Divisors across the accepted roots' range [2^-384, 2^500), quotients
    across [2^-400, 2^400), both signs; one sample in four an all-ones
    divisor over a power of two."""
    out = []
    while len(out) < n:
        b = math.ldexp(mantissa(rng), rng.randrange(-384, 500))
        q = math.ldexp(mantissa(rng), rng.randrange(-399, 399))
        a = float(Fraction(q) * Fraction(b))
        if rng.random() < 0.25:
            b = math.ldexp(2.0 - 2.0 ** -52, rng.randrange(-384, 499))
            a = math.ldexp(1.0, rng.randrange(-384, 399) + math.frexp(b)[1])
        if rng.random() < 0.5:
            a = -a
        if a != 0.0 and math.isfinite(a):
            out.append((a, b))
    return out


def root_samples(n, rng):
    out = []
    for _ in range(n):
        e = rng.randrange(-767, 1000)
        out.append(math.ldexp(mantissa(rng), e))
    return out


def rn_reciprocal(b):
    return float(1 / Fraction(b))


def test_final_correction_is_correctly_rounded():
    rng = random.Random(20261204)
    bad = 0
    for a, b in quotient_samples(4000, rng):
        q = div(a, b, rn_reciprocal(b))
        bad += q != a / b
    assert bad == 0


def test_refinement_keeps_and_reaches_the_reciprocal():
    rng = random.Random(5)
    keep = near = hard = 0
    for _, b in quotient_samples(3000, rng):
        y = rn_reciprocal(b)
        keep += rcp(b, y) != y
        all_ones = (math.frexp(b)[0] * 2) == 2.0 - 2.0 ** -52
        if not all_ones:
            near += rcp(b, seed(y, rng)) != y
        else:
            hard += rcp(b, math.nextafter(y, 0.0)) == y
    assert keep == 0 and near == 0
    assert hard == 0          # one unit low stays one unit low for all-ones divisors


def test_shared_reciprocal_serves_every_numerator():
    """One refined reciprocal, several numerators (a triangle gradient divides
    twice by each of its three roots)."""
    rng = random.Random(7)
    bad = 0
    for _, b in quotient_samples(1000, rng):
        y = rcp(b, rn_reciprocal(b))
        for _ in range(4):
            q = math.ldexp(mantissa(rng), rng.randrange(-60, 60))
            a = float(Fraction(q) * Fraction(b)) * (1 + 2.0 ** -40 * rng.random())
            bad += div(a, b, y) != a / b
    assert bad == 0


def test_roots_are_correctly_rounded():
    rng = random.Random(11)
    bad = 0
    for x in root_samples(4000, rng):
        y = seed(1.0 / math.sqrt(x), rng)
        bad += sqrt_in_range(x, y) != math.sqrt(x)
    assert bad == 0


def test_the_samples_reject_insufficient_refinements():
    """Planted mistakes: a coarse reciprocal seed; one Newton step fewer from
    a coarse seed; a coarse root seed without the corrections.  The same
    samples must catch each."""
    rng = random.Random(20261204)
    coarse = short = 0
    for a, b in quotient_samples(1000, rng):
        coarse += div(a, b, rcp(b, seed(1.0 / b, rng, coarse=True))) != a / b
        short += div(a, b, rcp(b, seed(1.0 / b, rng, coarse=True), steps=1)) != a / b
    rng = random.Random(11)
    bad_r = 0
    for x in root_samples(1000, rng):
        bad_r += sqrt_in_range(x, seed(1.0 / math.sqrt(x), rng, coarse=True), corrections=0) != math.sqrt(x)
    assert coarse > 10 and short > 100 and bad_r > 100
