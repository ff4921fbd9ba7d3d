"""The device's FP64 quotients and square roots against IEEE arithmetic
(numpy on the host).

Every kernel divides with one routine (and the shared-divisor form built on
the same reciprocal); the reference divides with the host's IEEE division,
so the two must agree bit for bit.  The cases are the ones that break a
reciprocal left one unit away from the nearest: divisors 1 - k 2^-53 with
small odd k, whose reciprocals lie next to a midpoint, paired with
numerators that put the quotient next to a midpoint too; divisors with an
all-ones significand; and random operands over the whole range.
"""
import numpy as np
import pytest

from sams_amd import lare

pytestmark = pytest.mark.gpu


def _bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def _agree(a, d, mode):
    q = lare.div_check(a, d, mode)
    want = a / d
    bad = (_bits(q) != _bits(want)) & ~(np.isnan(q) & np.isnan(want))
    idx = np.flatnonzero(bad)
    assert idx.size == 0, (
        f"mode {mode}: {idx.size} of {a.size} quotients differ, first "
        f"{a[idx[0]].hex()} / {d[idx[0]].hex()}: device {q[idx[0]].hex()} host {want[idx[0]].hex()}")


def _near_midpoint(rng, n, kmax):
    """a / d with d = 1 - k 2^-53 (k odd, < kmax) and a = m 2^-52 chosen so
    that m k = 2^52 + s (mod 2^53), |s| <= 64: the exact quotient lies within
    (|s| + k^2 2^-53) 2^-53 units of a midpoint."""
    a, d = [], []
    M = 1 << 53
    while len(a) < n:
        k = int(rng.integers(0, kmax // 2)) * 2 + 1
        s = int(rng.integers(-64, 65))
        m = ((1 << 52) + s) * pow(k, -1, M) % M
        if m < (1 << 52):
            continue
        ea, ed = (int(v) for v in rng.integers(-20, 21, 2))
        a.append(np.ldexp(float(m), -52 + ea) * (1 if rng.integers(2) else -1))
        d.append(np.ldexp(1.0 - np.ldexp(float(k), -53), ed))
    return np.array(a), np.array(d)


@pytest.mark.parametrize("mode", [0, 1])
def test_quotient_found_in_a_wind_tunnel_step(mode):
    # rho / (1 + dv) of one cell of the Emery wind tunnel's fourth step
    a = np.array([1.3999999999999992])
    d = np.array([1.0 - 5 * 2.0 ** -53])
    assert (a / d)[0] == 1.4000000000000001
    _agree(a, d, mode)


@pytest.mark.parametrize("mode", [0, 1])
def test_reciprocals_of_divisors_just_below_one(mode):
    k = np.arange(1, 1 << 16, dtype=np.float64)
    d = 1.0 - k * 2.0 ** -53
    for e in (0, -7, 12):
        _agree(np.ones_like(d), np.ldexp(d, e), mode)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("kmax", [1 << 8, 1 << 12, 1 << 16])
def test_quotients_next_to_midpoints(mode, kmax):
    a, d = _near_midpoint(np.random.default_rng(kmax), 1 << 16, kmax)
    _agree(a, d, mode)


@pytest.mark.parametrize("mode", [0, 1])
def test_all_ones_significands(mode):
    rng = np.random.default_rng(7)
    n = 1 << 14
    d = np.ldexp(np.full(n, 2.0 - 2.0 ** -52), rng.integers(-100, 101, n))
    a = np.ldexp(np.where(np.arange(n) % 2 == 0, 1.0, rng.uniform(1.0, 2.0, n)), rng.integers(-100, 101, n))
    d[1::4] = -d[1::4]
    a[2::4] = -a[2::4]
    _agree(a, d, mode)


@pytest.mark.parametrize("mode", [0, 1])
def test_all_ones_significands_at_the_exponent_limits(mode):
    """Divisors with an all-ones significand at the top and the bottom of
    the exponent range, subnormal ones included (whose significand field is
    all ones, or all ones after normalisation), of both signs, against
    numerators that keep the quotient normal."""
    rng = np.random.default_rng(19)
    sub = np.array([0x000FFFFFFFFFFFFF, 0x0007FFFFFFFFFFFF, 0x0003FFFFFFFFFFFF, 0x00000000FFFFFFFF,
                    0x00000000000001FF, 0x0000000000000001], dtype=np.uint64).view(np.float64)
    ones = 2.0 - 2.0 ** -52
    d = np.concatenate([sub, np.ldexp(ones, np.arange(990, 1024)), np.ldexp(ones, np.arange(-1022, -990))])
    d = np.concatenate([d, -d])
    m = 256
    ed = np.frexp(d)[1]
    ea = np.clip(ed[:, None] + rng.integers(-60, 61, (d.size, m)), -1060, 1020)
    a = np.ldexp(np.where(rng.integers(0, 2, (d.size, m)) == 0, 1.0, rng.uniform(1.0, 2.0, (d.size, m))), ea)
    a[:, 1::2] = -a[:, 1::2]
    dd = np.repeat(d, m)
    with np.errstate(all="ignore"):
        _agree(a.ravel(), dd, mode)


@pytest.mark.parametrize("mode", [0, 1])
def test_random_operands(mode):
    rng = np.random.default_rng(11)
    n = 1 << 20
    a = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-60, 61, n)) * rng.choice([-1.0, 1.0], n)
    d = np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-60, 61, n)) * rng.choice([-1.0, 1.0], n)
    _agree(a, d, mode)


@pytest.mark.parametrize("mode", [0, 1])
def test_special_and_extreme_operands(mode):
    rng = np.random.default_rng(13)
    a = rng.integers(0, 1 << 63, 1 << 18, dtype=np.uint64).view(np.float64).copy()
    d = rng.integers(0, 1 << 63, 1 << 18, dtype=np.uint64).view(np.float64).copy()
    a[1::2] = -a[1::2]
    special = [(0.0, 0.0), (1.0, 0.0), (-0.0, 3.0), (0.0, -3.0), (np.inf, np.inf), (np.inf, 2.0),
               (2.0, np.inf), (np.nan, 1.0), (1e-310, 3.0), (1.0, 1e-310), (1e300, 1e-300),
               (1e-300, 1e300), (5e-324, 0.5), (np.finfo(float).max, 0.5)]
    for i, (x, y) in enumerate(special):
        a[i], d[i] = x, y
    with np.errstate(all="ignore"):
        _agree(a, d, mode)


def _near_root_midpoints(rng, n):
    """x in [1, 4) (scaled by even powers of two) whose square root lies
    within 1/256 unit of a midpoint, found with the host's 64-bit-significand
    long double."""
    out = []
    while len(out) < n:
        x = rng.uniform(1.0, 4.0, 1 << 16)
        r = np.sqrt(x.astype(np.longdouble))
        lo = r.astype(np.float64)
        up = np.nextafter(lo, np.where(r > lo, np.inf, -np.inf))
        mid = (lo.astype(np.longdouble) + up.astype(np.longdouble)) / 2
        dist = np.abs(r - mid) / np.abs(up.astype(np.longdouble) - lo.astype(np.longdouble))
        out.extend(x[dist < 1.0 / 256].tolist())
    x = np.array(out[:n])
    return np.ldexp(x, 2 * rng.integers(-100, 101, n))


def test_square_roots():
    """sqrt as the kernels compute it (hipcc's sequence) against the host's."""
    rng = np.random.default_rng(17)
    assert np.finfo(np.longdouble).nmant >= 63, "needs the x86 long double"
    x = np.concatenate([
        _near_root_midpoints(rng, 1 << 15),
        np.ldexp(rng.uniform(1.0, 4.0, 1 << 18), 2 * rng.integers(-200, 201, 1 << 18)),
        rng.integers(0, 0x7ff0000000000000, 1 << 18, dtype=np.uint64).view(np.float64),
        np.array([0.0, -0.0, np.inf, 5e-324, 1e-310, np.finfo(float).max, 2.0, 3.0]),
    ])
    q = lare.div_check(x, np.ones_like(x), 2)
    want = np.sqrt(x)
    bad = np.flatnonzero(_bits(q) != _bits(want))
    assert bad.size == 0, f"{bad.size} of {x.size} roots differ, first sqrt({x[bad[0]].hex()}): device {q[bad[0]].hex()} host {want[bad[0]].hex()}"
