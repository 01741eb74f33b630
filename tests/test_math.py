"""The shared IEEE numerics (mc_math.h) against the host libm, which is
independent of it: sin/cos must be within 1 ulp of libm's binary64 result
rounded to binary32 everywhere, and in practice equal to it.  On the GPU: the
ray fast path's sqrt / reciprocal / quotient cores against the full IEEE
operations."""
import math

import numpy as np
import pytest

import oracle


def _ulp(a, b):
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


def _ref(x):
    s = np.array([math.sin(float(v)) for v in x], np.float64).astype(np.float32)
    c = np.array([math.cos(float(v)) for v in x], np.float64).astype(np.float32)
    return s, c


def test_sincos_random_bit_patterns():
    rng = np.random.default_rng(0)
    x = rng.integers(0, 2**32, 200_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    x = x[np.isfinite(x)]
    s, c = oracle.sincosf(x)
    rs, rc = _ref(x)
    assert _ulp(s, rs).max() <= 1 and _ulp(c, rc).max() <= 1
    assert (s == rs).mean() > 0.9999 and (c == rc).mean() > 0.9999


def test_sincos_near_multiples_of_half_pi():
    """Where argument reduction cancels: floats next to k*pi/2, |k| <= 2^16,
    plus a sweep of huge arguments (Payne-Hanek path)."""
    k = np.arange(-(1 << 16), 1 << 16, 7, dtype=np.float64)
    base = (k * (math.pi / 2)).astype(np.float32)
    x = np.concatenate([base, np.nextafter(base, np.float32(np.inf)),
                        np.nextafter(base, np.float32(-np.inf)),
                        np.float32(2.0) ** np.arange(-30, 128, dtype=np.float32) * np.float32(1.2345)])
    x = x[np.isfinite(x)]
    s, c = oracle.sincosf(x)
    rs, rc = _ref(x)
    assert _ulp(s, rs).max() <= 1 and _ulp(c, rc).max() <= 1


def test_sincos_special_values():
    x = np.float32([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4028235e38])
    s, c = oracle.sincosf(x)
    assert s[0] == 0 and not np.signbit(s[0]) and c[0] == 1
    assert s[1] == 0 and np.signbit(s[1]) and c[1] == 1
    assert np.isnan(s[2:5]).all() and np.isnan(c[2:5]).all()
    assert s[5] == np.float32(1e-45) and s[6] == np.float32(-1e-45)
    rs, rc = _ref(x[7:])
    assert _ulp(s[7:], rs).max() <= 1 and _ulp(c[7:], rc).max() <= 1


def test_half_fov_tangent():
    for fov in (30.0, 60.0, 70.0, 90.0, 110.0):
        half = np.float32(np.float32(np.float32(fov) * np.float32(3.14159265)) / np.float32(180.0)) \
            * np.float32(0.5)
        ref = np.float32(math.tan(float(half)))
        got = np.float32(oracle.half_fov_tan(fov))
        assert _ulp(np.array([got]), np.array([ref]))[0] <= 1, fov


@pytest.mark.gpu
@pytest.mark.parametrize("mode,ranges", [
    # every float with |x| in [2^-90, 2^90] (both signs)
    (0, [(0x12800000, 0x6C800000 - 0x12800000), (0x92800000, 0x6C800000 - 0x12800000)]),
    # every float in [2^-60, 2^60]
    (1, [(0x21800000, 0x5D800000 - 0x21800000)]),
    # 2^32 hashed quotient pairs
    (2, [(0, 0xFFFFFFFF)]),
])
def test_ray_setup_cores_equal_ieee(mode, ranges):
    """The cores of the ray fast path (sqrt, reciprocal, quotient without
    the IEEE sequences' scaling and fix-ups) return the same bits as the
    full IEEE operations over their whole guarded operand range (reciprocal,
    sqrt: exhaustive) or on 2^32 hashed pairs (quotient)."""
    import ctypes
    import torch
    from minecu import _lib
    L = _lib.load()
    out = torch.zeros(3, dtype=torch.int32, device="cuda")
    for begin, count in ranges:
        _lib.check(L.mc_debug_fast_math(mode, begin, count, ctypes.c_void_p(out.data_ptr()),
                                        torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    checked, bad, bad_in = (int(v) & 0xFFFFFFFF for v in out.cpu().tolist())
    assert checked > 10 ** 9
    assert bad == 0, f"{bad} mismatches, e.g. input bits {bad_in:#010x}"
