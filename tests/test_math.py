"""The shared scalar primitives (include/clim_math.h) against the C library
(numpy), published known-answer vectors and exact integer arithmetic.

These primitives are the only code the FP32 oracle shares with the kernels,
so each one is pinned here with an explicit error bound."""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def L():
    return O.lib()


def _vec(L, name, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    getattr(L, name)(x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.c_long(x.size))
    return out


def test_atan2pi(L):
    rng = np.random.default_rng(1)
    y = np.concatenate([rng.normal(size=200000), rng.uniform(0, 1e-3, 1000), [0, 0, 1, -1, 0, 1e-30, 3]]).astype(np.float32)
    x = np.concatenate([rng.normal(size=200000), rng.uniform(0, 1e-3, 1000), [0, 1, 0, 0, -1, 1e-30, 3]]).astype(np.float32)
    yo = np.ascontiguousarray(y)
    out = np.empty_like(yo)
    L.om_atan2pi(yo.ctypes.data_as(C.c_void_p), np.ascontiguousarray(x).ctypes.data_as(C.c_void_p),
                 out.ctypes.data_as(C.c_void_p), C.c_long(y.size))
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64)) / np.pi
    err = np.abs(out - ref)
    assert err.max() < 2.5e-7, err.max()
    # exact special values
    assert out[-7] == 0.0 and out[-6] == 0.0 and out[-5] == 0.5 and out[-4] == -0.5 and out[-3] == 1.0
    assert out[-1] == 0.25


def test_exp_log_tanh_cos(L):
    rng = np.random.default_rng(2)
    x = rng.uniform(-80, 80, 200000).astype(np.float32)
    e = _vec(L, "om_exp", x)
    rel = np.abs(e / np.exp(x.astype(np.float64)) - 1)
    assert rel.max() < 4e-7, rel.max()
    pos = np.concatenate([rng.uniform(1e-30, 10, 100000), 10.0 ** rng.uniform(-40, 30, 100000)]).astype(np.float32)
    lg = _vec(L, "om_log", pos)
    err = np.abs(lg - np.log(pos.astype(np.float64)))
    assert (err / np.maximum(np.abs(np.log(pos.astype(np.float64))), 1.0)).max() < 3e-7
    t = rng.uniform(-12, 12, 200000).astype(np.float32)
    th = _vec(L, "om_tanh", t)
    assert np.abs(th - np.tanh(t.astype(np.float64))).max() < 3e-7
    u = rng.uniform(-3, 3, 200000).astype(np.float32)
    c = _vec(L, "om_cos2pi", u)
    assert np.abs(c - np.cos(2 * np.pi * u.astype(np.float64))).max() < 4e-7
    assert _vec(L, "om_exp", np.array([0.0], np.float32))[0] == 1.0
    assert _vec(L, "om_log", np.array([1.0], np.float32))[0] == 0.0


def _vec2(L, name, x, y):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    getattr(L, name)(x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                     C.c_long(x.size))
    return out


def test_pow_accuracy_and_c99_special_values(L):
    rng = np.random.default_rng(5)
    x = (10.0 ** rng.uniform(-6, 3, 300000)).astype(np.float32)
    y = rng.uniform(-3, 3, 300000).astype(np.float32)
    ref = np.power(x.astype(np.float64), y.astype(np.float64))
    keep = np.abs(y.astype(np.float64) * np.log(x.astype(np.float64))) <= 20
    got = _vec2(L, "om_pow", x, y)
    rel = np.abs(got[keep] / ref[keep] - 1)
    assert rel.max() < 4e-6, rel.max()
    small = np.abs(y.astype(np.float64) * np.log(x.astype(np.float64))) <= 4
    assert np.abs(got[small] / ref[small] - 1).max() < 1e-6
    # every pairing of the special operands against C's pow (numpy float64):
    # NaN-ness, sign and value (exact for the special results)
    sv = np.array([0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, np.inf, -np.inf, np.nan, 3.0, -3.0, 2.5, -2.5,
                   1e-3, 16777216.0, -16777217.0], np.float32)
    xs, ys = np.meshgrid(sv, sv, indexing="ij")
    xs, ys = xs.ravel(), ys.ravel()
    with np.errstate(all="ignore"):
        ref = np.power(xs.astype(np.float64), ys.astype(np.float64)).astype(np.float32)
    got = _vec2(L, "om_pow", xs, ys)
    for a, b, g, r in zip(xs, ys, got, ref):
        if np.isnan(r):
            assert np.isnan(g), (a, b, g)
            continue
        assert not np.isnan(g) and np.signbit(g) == np.signbit(r), (a, b, g, r)
        if np.isinf(r) or r == 0 or abs(np.log(abs(float(a))) * float(b)) < 1e-30:
            assert g == r, (a, b, g, r)
        else:
            assert abs(g / r - 1) < 4e-6, (a, b, g, r)


def test_smoothstep_opencl_semantics(L):
    x = np.array([-1, 0.0, 0.25, 0.5, 1.0, 2.0, np.nan], np.float32)
    out = np.empty_like(x)
    L.om_smoothstep(C.c_float(0.0), C.c_float(1.0), x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                    C.c_long(x.size))
    t = np.clip(x[:-1].astype(np.float64), 0, 1)
    np.testing.assert_allclose(out[:-1], t * t * (3 - 2 * t), rtol=0, atol=1e-7)
    assert out[-1] == 0.0  # clamp(NaN) = fmin(fmax(NaN, 0), 1) = 0
    # zero-width edges: x > e gives 1, x == e gives 0 (0/0 -> NaN -> 0)
    x2 = np.array([0.0, 1.0], np.float32)
    L.om_smoothstep(C.c_float(0.0), C.c_float(0.0), x2.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                    C.c_long(2))
    assert out[0] == 0.0 and out[1] == 1.0


def test_philox_known_answers(L):
    """Random123 known-answer vectors for philox4x32-10."""
    out = (C.c_uint32 * 4)()

    def run(ctr, k0, k1):
        L.om_philox((C.c_uint32 * 4)(*ctr), C.c_uint32(k0), C.c_uint32(k1), out)
        return list(out)

    assert run([0, 0, 0, 0], 0, 0) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert run([0xffffffff] * 4, 0xffffffff, 0xffffffff) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert run([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], 0xa4093822, 0x299f31d0) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_fixed_point_is_exact_truncation(L):
    rng = np.random.default_rng(3)
    t = np.concatenate([rng.uniform(0, 0.5, 2000), rng.uniform(0, 1e-3, 2000) ** 2,
                        (rng.uniform(0, 0.5, 1000).astype(np.float32).astype(np.float64)) ** 2,
                        [0.0, 0.5, 2.0 ** -90, 2.0 ** -91, 3.0e10]])
    out = np.empty(2 * t.size, np.uint64)
    L.om_fx_from_double(t.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.c_long(t.size))
    for i in range(t.size):
        exact = int(Fraction(float(t[i])) * (1 << 90))  # floor for t >= 0
        got = (int(out[2 * i]) << 64) | int(out[2 * i + 1])
        assert got == exact, (t[i], got, exact)
    # the sum is independent of order
    L.om_fx_sum.restype = C.c_double
    hi1, lo1, hi2, lo2 = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
    a = np.ascontiguousarray(t)
    b = np.ascontiguousarray(t[rng.permutation(t.size)])
    s1 = L.om_fx_sum(a.ctypes.data_as(C.c_void_p), C.c_long(a.size), C.byref(hi1), C.byref(lo1))
    s2 = L.om_fx_sum(b.ctypes.data_as(C.c_void_p), C.c_long(b.size), C.byref(hi2), C.byref(lo2))
    assert (hi1.value, lo1.value) == (hi2.value, lo2.value) and s1 == s2
    assert abs(s1 - float(np.sum(np.sort(t)))) <= 1e-12 * s1


def test_q32_terms_are_floors_on_the_grid(L):
    """mq = floor(m 2^32) and vq = floor(v 2^41) against exact rationals, over
    the in-range domain and its ends (mq <= 2^31, vq < 2^32)."""
    from fractions import Fraction
    rng = np.random.default_rng(9)
    m_min = np.float32(0.001)
    if m_min < 0.001:
        m_min = np.nextafter(m_min, np.float32(1))
    v_max = np.float32(0.001)
    if v_max > 0.001:
        v_max = np.nextafter(v_max, np.float32(0))
    m = np.concatenate([rng.uniform(0.001, 0.5, 3000), 10.0 ** rng.uniform(-3, np.log10(0.5), 3000)])
    m = np.concatenate([np.clip(m, m_min, 0.5), [m_min, 0.5, np.nextafter(np.float32(0.5), np.float32(0)), 0.25]])
    v = np.concatenate([rng.uniform(0, 1e-3, 3000), 10.0 ** rng.uniform(-40, -3, 3000)])
    v = np.concatenate([np.clip(v, 0, v_max), [0.0, v_max, 1e-45, 2.0 ** -41]])
    m, v = m.astype(np.float32), v.astype(np.float32)
    out = np.empty(2 * m.size, np.uint32)
    L.om_q32_terms(m.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p),
                   C.c_long(m.size))
    o = out.reshape(-1, 2)
    for i in range(m.size):
        assert int(o[i, 0]) == int(Fraction(float(m[i])) * 2 ** 32), m[i]
        assert int(o[i, 1]) == int(Fraction(float(v[i])) * 2 ** 41), v[i]
    assert int(o[np.flatnonzero(m == 0.5)[0], 0]) == 2 ** 31
    assert int(o[np.flatnonzero(v == v_max)[0], 1]) < 2 ** 32


def test_fit_q_solve_and_moment_residuals(L):
    """clim_fit_solve_q against least squares on the grid values, and
    clim_fit_rss_q (moment expansion) against the residuals summed exactly."""
    from fractions import Fraction
    rng = np.random.default_rng(10)
    for trial in range(6):
        n = int(rng.integers(2, 3000))
        mq = rng.integers(2 ** 22, 2 ** 31, n, dtype=np.int64)
        slope, icpt = rng.uniform(1e-4, 5e-3), rng.uniform(1e-6, 5e-5)
        vf = icpt + slope * mq / 2.0 ** 32 + rng.normal(0, 10.0 ** rng.uniform(-7, -5), n)
        vq = np.clip(np.floor(vf * 2.0 ** 41), 0, 2 ** 32 - 1).astype(np.int64)
        if trial == 5:  # samples exactly on a line: residuals vanish
            vq = 1000 + 3 * mq // 2 ** 11
        mi, vi = [int(x) for x in mq], [int(x) for x in vq]
        sums = [sum(mi), sum(vi), sum(x * x for x in mi), sum(x * y for x, y in zip(mi, vi)), sum(y * y for y in vi)]
        flat = np.array([w for t in sums for w in (t >> 64, t & (2 ** 64 - 1))], np.uint64)
        out = np.empty(3, np.float64)
        L.om_fit_q(flat.ctypes.data_as(C.c_void_p), C.c_ulonglong(n), out.ctypes.data_as(C.c_void_p))
        A, B, rss = out
        mg, vg = mq / 2.0 ** 32, vq / 2.0 ** 41
        Bl, Al = np.polyfit(mg, vg, 1)
        if Bl > 1e-8 and Al > 1e-8:
            np.testing.assert_allclose([A, B], [Al, Bl], rtol=1e-8, atol=1e-14)
        exact = sum((Fraction(A) + Fraction(B) * Fraction(x, 2 ** 32) - Fraction(y, 2 ** 41)) ** 2
                    for x, y in zip(mi, vi))
        scale = sum(Fraction(y, 2 ** 41) ** 2 for y in vi)
        assert abs(Fraction(rss) - exact) <= Fraction(1, 10 ** 12) * scale + Fraction(1, 10 ** 30), trial


def test_tanh_across_the_clamp(L):
    """clim_tanhf (the exponential inlined in fused form, argument clamped at
    -20): within 3e-7 of tanh on samples across the clamp, odd, exactly +-1
    where (1 - e) / (1 + e) rounds to 1, and the special values."""
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-12, 12, 200000), rng.uniform(-200, 200, 50000),
                        10.0 ** rng.uniform(-40, 38, 50000) * rng.choice([-1, 1], 50000),
                        [0.0, -0.0, 10.0, -10.0, 9.999, 10.001, 30.0, -30.0, 1e30]]).astype(np.float32)
    th = _vec(L, "om_tanh", x)
    assert np.abs(th - np.tanh(x.astype(np.float64))).max() < 3e-7
    assert np.array_equal(_vec(L, "om_tanh", -x).view(np.uint32), (-th).view(np.uint32))
    big = np.abs(x) >= 9.1  # e < 2^-26: the quotient rounds to 1
    assert np.all(np.abs(th[big]) == 1.0)
    assert np.signbit(_vec(L, "om_tanh", np.array([-0.0], np.float32))[0])
    assert np.isnan(_vec(L, "om_tanh", np.array([np.nan], np.float32))[0])
    assert _vec(L, "om_tanh", np.array([np.inf, -np.inf], np.float32)).tolist() == [1.0, -1.0]


@pytest.mark.parametrize("e0,e1", [(0.9, 2.0), (0.75, 2.0), (0.0, 0.3)])
def test_smoothstep_rc_equals_division(L, e0, e1):
    """The tone map's smoothsteps with the division by the edge span as a
    reciprocal multiplication and one fused correction equal the literal
    form: over every float between the edges (where the quotient is in
    [0, 1]; outside, both forms clamp to the same end since the corrected
    quotient is within an ulp of the true one), a sample of all other
    patterns and the special values (infinities, NaN, zeros)."""
    L.om_smoothstep_rc_mismatches.restype = C.c_long
    L.om_smoothstep_rc_mismatches.argtypes = [C.c_float, C.c_float, C.c_uint32, C.c_uint32]
    f0, f1 = np.float32(e0), np.float32(e1)
    lo, hi = int(f0.view(np.uint32)), int(f1.view(np.uint32)) + 1
    assert L.om_smoothstep_rc_mismatches(f0, f1, lo, hi) == 0
    for start in range(0, 2 ** 32, 2 ** 26):  # every 64th block of 2^20 patterns
        assert L.om_smoothstep_rc_mismatches(f0, f1, start, start + 2 ** 20) == 0, hex(start)
    for v in (0x7F800000, 0xFF800000, 0x7FC00000, 0x00000000, 0x80000000, 0x7F7FFFFF, 0xFF7FFFFF):
        assert L.om_smoothstep_rc_mismatches(f0, f1, v, v + 1) == 0, hex(v)


def test_div65535_equals_division_for_every_uint16(L):
    """The kernels' scaleRawData quotient (reciprocal + one fused correction)
    is bit-identical to the IEEE division for all 65536 inputs."""
    out = np.zeros(65536, np.uint8)
    L.om_div65535_check(out.ctypes.data_as(C.c_void_p))
    assert out.all(), np.flatnonzero(out == 0)[:10]
    # and against numpy's float32 division
    v = np.arange(65536, dtype=np.float32)
    assert np.array_equal(v / np.float32(65535), (v / np.float32(65535)).astype(np.float32))


def test_div25_equals_division(L):
    """clim_div25 (reciprocal + fused correction) against x / 25.0f.  Division
    by a constant commutes with scaling by 2^k unless the quotient is
    subnormal (the correction term is then a multiple of a normal ulp, so
    exact), so every float of one ordinary binade plus every float below
    2^-121 (quotient subnormal) decide it, both signs; the build also checked
    all 2^32 patterns once."""
    L.om_div25_mismatches.restype = C.c_long
    L.om_div25nz_mismatches.restype = C.c_long
    ranges = [(0x3F800000, 0x40000000), (0x00000000, 0x03000000)]
    for lo, hi in ranges:
        for sign in (0, 0x80000000):
            assert L.om_div25_mismatches(C.c_uint32(lo | sign), C.c_uint32(hi | sign)) == 0, (hex(lo), sign)
            # the form without the zero test (window statistics, no -0 operand)
            assert L.om_div25nz_mismatches(C.c_uint32(lo | sign), C.c_uint32(hi | sign)) == 0, (hex(lo), sign)
    # ... which differs at -0 alone: it gives +0 there
    L.om_div25_nz.restype = C.c_float
    L.om_div25_nz.argtypes = [C.c_float]
    z = L.om_div25_nz(-0.0)
    assert z == 0.0 and np.signbit(np.float32(z)) == False


def test_div_count_equals_division(L):
    """clim_div_count (the bilateral's sum / count, n = 0..25) against x / n.
    As for clim_div25, division by a constant commutes with scaling by 2^k
    unless the quotient is subnormal, so one ordinary binade, every float
    below 2^-121 (quotients near and below the subnormal range) and the
    specials decide it, both signs; the build also checked all 2^32 patterns
    for every n.  A planted mistake (the form without its subnormal branch)
    is caught by the same ranges."""
    L.om_div_count_mismatches.restype = C.c_long
    L.om_div_count_mismatches.argtypes = [C.c_int, C.c_uint32, C.c_uint32]
    ranges = [(0x3F800000, 0x40000000), (0x00000000, 0x03000000), (0x7F800000, 0x80000000)]
    for n in range(26):
        for lo, hi in ranges:
            for sign in (0, 0x80000000):
                bad = L.om_div_count_mismatches(n, lo | sign, (hi - 1) | sign)
                assert bad == 0, (n, hex(lo), sign, bad)

    # the planted mistake: without the subnormal branch the low range fails
    L.om_div_count_nobranch_mismatches.restype = C.c_long
    L.om_div_count_nobranch_mismatches.argtypes = [C.c_int, C.c_uint32, C.c_uint32]
    assert L.om_div_count_nobranch_mismatches(6, 0x00000000, 0x02FFFFFF) > 0
    assert L.om_div_count_nobranch_mismatches(6, 0x3F800000, 0x3FFFFFFF) == 0


def test_scaled_step_equals_the_comparison(L):
    """The bilateral's weight as clamp(fma(l, -H, e0 H), 0, 1) (one GPU
    instruction) against l < e0: every l within 2^22 ulps of e0 on either
    side, every l of the positive binades from 2^-30 up to 2^8 and the
    specials, for the squared edges the kernels meet (2 without a learned
    map, 2 m^2 over the map's range m in [2^-1.5, 2^1.5], powers of two and
    their neighbours), and the smallest and largest edges it admits."""
    L.om_step_scaled_mismatches.restype = C.c_long
    L.om_step_scaled_mismatches.argtypes = [C.c_float, C.c_uint32, C.c_uint32]
    f = np.float32
    edges = [f(2.0), f(2.0) * f(2.0 ** -1.5) ** 2, f(2.0) * f(2.0 ** 1.5) ** 2, f(0.25), f(16.0),
             np.nextafter(f(2.0), f(0)), np.nextafter(f(2.0), f(4)), f(2.0 ** -77), f(3.0e38)]
    m = np.float32(np.random.default_rng(4).uniform(2.0 ** -1.5, 2.0 ** 1.5, 6))
    edges += [f(f(1.41421356) * f(1.41421356)) * mm * mm for mm in m]
    for e0 in edges:
        b = int(np.array(e0, np.float32).view(np.uint32))
        assert L.om_step_scaled_mismatches(e0, max(b - (1 << 22), 0), min(b + (1 << 22), 0x7F800000)) == 0, e0
        assert L.om_step_scaled_mismatches(e0, 0x7F800000, 0x80000000) == 0, e0  # inf, NaNs
        assert L.om_step_scaled_mismatches(e0, 0xFF800000, 0xFFFFFFFF) == 0, e0
        assert L.om_step_scaled_mismatches(e0, 0x80000000, 0x80400000) == 0, e0  # -0 and negatives
    for e0 in edges[:7]:
        assert L.om_step_scaled_mismatches(e0, 0x30800000, 0x43800000) == 0, e0


def test_window_sums_are_never_negative_zero():
    """The premise of clim_div25_nz in the window statistics: under
    round-to-nearest, +0 + (-0) = +0 and x + (-x) = +0, so a float sum
    that starts from +0 is never -0, nor is a square or the difference of
    two such values; checked on sums of signed zeros and cancelling terms."""
    f = np.float32
    vals = [f(0.0), f(-0.0), f(1.5), f(-1.5), f(2.0 ** -149), f(-(2.0 ** -149))]
    rng = np.random.default_rng(5)
    for _ in range(2000):
        terms = rng.choice(len(vals), 5)
        s = f(0.0)
        for t in terms:
            s = f(s + vals[t])
        assert not (s == 0 and np.signbit(s)), terms
        sq = f(s * s)
        assert not np.signbit(sq)
        d = f(sq - sq)
        assert not (d == 0 and np.signbit(d))


def _median9_network(q):
    """The forgetful-selection network of fast_median3x3 (cl/demosaic.cl:384-413)
    with the comparison-select min/max of the reference's macros."""
    def s(a, b):
        return (a if a < b else b), (a if a > b else b)
    a0, a1, a2, a3, a4, a5 = q[:6]
    a0, a1 = s(a0, a1); a2, a3 = s(a2, a3); a4, a5 = s(a4, a5); a0, a2 = s(a0, a2)
    a1, a3 = s(a1, a3); a0, a4 = s(a0, a4); a3, a5 = s(a3, a5)
    a0 = q[6]
    a0, a1 = s(a0, a1); a2, a3 = s(a2, a3); a0, a2 = s(a0, a2); a1, a3 = s(a1, a3); a0, a4 = s(a0, a4)
    a3, a4 = s(a3, a4)
    a0 = q[7]
    a0, a1 = s(a0, a1); a2, a3 = s(a2, a3); a0, a2 = s(a0, a2); a1, a3 = s(a1, a3)
    a0 = q[8]
    a0, a1 = s(a0, a1); a0, a2 = s(a0, a2); a1, a2 = s(a1, a2)
    return a1


def _med3(a, b, c):
    return sorted((a, b, c))[1]


def test_despeckle_nan_free_forms():
    """The despeckle's NaN-free form (the fused red/blue + despeckle kernel takes
    it for tiles without NaN): the median of nine as med3(max of row minima,
    median of row medians, min of row maxima) over rows of three, and the
    luma trackers as a running max/min plus a median of three, give the same
    values as the reference's selection network and comparison trackers.
    Exhaustive over {0..3}^9 (every order and tie pattern of up to four
    levels) and over {-1, 0, 0.5, 1, 70000}^7 for the trackers (values on
    both sides of the 0 and 65535 starting points), plus random reals."""
    import itertools

    def fast_median(q):
        rows = [q[0:3], q[3:6], q[6:9]]
        return _med3(max(min(r) for r in rows), _med3(*[_med3(*r) for r in rows]), min(max(r) for r in rows))

    for q in itertools.product(range(4), repeat=9):
        m = sorted(q)[4]
        assert _median9_network(q) == m and fast_median(q) == m, q

    def trackers_ref(vs):
        fM = sM = 0.0
        fm = sm = 65535.0
        for v in vs:
            sM = v if (v <= fM and v > sM) else sM
            sM = fM if v > fM else sM
            fM = v if v > fM else fM
            sm = v if (v >= fm and v < sm) else sm
            sm = fm if v < fm else sm
            fm = v if v < fm else fm
        return fM, sM, fm, sm

    def trackers_fast(vs):
        fM = sM = 0.0
        fm = sm = 65535.0
        for v in vs:
            sM, fM = _med3(sM, v, fM), max(fM, v)
            sm, fm = _med3(fm, v, sm), min(fm, v)
        return fM, sM, fm, sm

    for q in itertools.product([-1.0, 0.0, 0.5, 1.0, 70000.0], repeat=7):
        assert trackers_ref(q) == trackers_fast(q), q
    rng = np.random.default_rng(9)
    for _ in range(20000):
        q = list(rng.normal(size=9) * 3)
        assert _median9_network(q) == fast_median(q) == sorted(q)[4]
        assert trackers_ref(q) == trackers_fast(q)


def test_rsqrt_accuracy_and_special_values(L):
    """clim_rsqrt (the denoise's 1/sigma): within 2.5 ulp of 1/sqrt(s) over the
    binade pair [1, 4), which decides every s in [2^-124, 2^124] (s -> 4s
    halves every intermediate exactly), within 4 ulp over the binades at both
    ends of the normal range; NaN for s < 0 and NaN, +inf at zero, 0 at +inf."""
    L.om_rsqrt_max_rel_err.restype = C.c_double
    L.om_rsqrt_max_rel_err.argtypes = [C.c_uint32, C.c_uint32]
    L.om_rsqrt.restype = C.c_float
    L.om_rsqrt.argtypes = [C.c_float]
    ulp = 2.0 ** -24
    assert L.om_rsqrt_max_rel_err(0x3F800000, 0x40800000) <= 2.5 * ulp
    for lo in (0x01800000, 0x20000000, 0x5F000000, 0x7D800000):
        assert L.om_rsqrt_max_rel_err(lo, lo + (1 << 20)) <= 2.5 * ulp, hex(lo)
    assert L.om_rsqrt_max_rel_err(0x00800000, 0x01800000) <= 4.0 * ulp
    assert L.om_rsqrt_max_rel_err(0x7E000000, 0x7F800000) <= 4.0 * ulp
    # the period: s and 4 s give results exactly a factor 2 apart
    for s in (1.0, 1.37, 2.9, 3.999):
        a, b = L.om_rsqrt(s), L.om_rsqrt(4.0 * s)
        assert np.float32(a) == np.float32(2.0) * np.float32(b)
    assert np.isnan(L.om_rsqrt(-1.0)) and np.isnan(L.om_rsqrt(float("nan")))
    assert L.om_rsqrt(0.0) == float("inf") and L.om_rsqrt(float("inf")) == 0.0
