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


def test_integer_fixed_point_conversions_match_double_path(L):
    rng = np.random.default_rng(4)
    f = np.concatenate([rng.uniform(0, 0.5, 20000), rng.uniform(0, 1e-3, 20000), 10.0 ** rng.uniform(-45, 10, 20000),
                        [0.0, -0.0, 1e-45, 1.17549435e-38, 0.5, 0.001]]).astype(np.float32)
    g = rng.permutation(f).astype(np.float32)
    out = np.empty(4 * f.size, np.uint64)
    L.om_fx_float_check(f.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.c_long(f.size))
    o = out.reshape(-1, 4)
    np.testing.assert_array_equal(o[:, :2], o[:, 2:])
    L.om_fx_product_check(f.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p),
                          out.ctypes.data_as(C.c_void_p), C.c_long(f.size))
    o = out.reshape(-1, 4)
    np.testing.assert_array_equal(o[:, :2], o[:, 2:])
    t = np.concatenate([rng.uniform(0, 1e-6, 20000) ** 2, 10.0 ** rng.uniform(-300, 10, 20000),
                        [0.0, -0.0, 5e-324, 2.0 ** -90, 2.0 ** -91, np.nan]])
    out = np.empty(4 * t.size, np.uint64)
    L.om_fx_double_bits_check(t.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.c_long(t.size))
    o = out.reshape(-1, 4)
    np.testing.assert_array_equal(o[:, :2], o[:, 2:])
