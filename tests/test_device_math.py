"""The correctly rounded helpers the pair kernels use instead of the
compiler's division and square root (physica_amd/csrc/device_math.hpp):

- div_rn(x, b, RN(1/b)) must equal x / b. The sequence is a multiply and four
  fused multiply-adds, so a power-of-two scaling of x scales every step
  exactly: checking every x of a binade checks every x whose quotient and
  remainders stay normal. The host check covers two binades below each
  support radius the scenes use (and its viscosity constant pi h^6) plus
  random divisors; the GPU check runs every float in [2^-100, h). Below
  2^-105 the remainders underflow and the sequence can miss (x = 0 is exact);
  the kernels' arguments are distances, at least 2^-74.5 when nonzero.
- sqrt_rn(x) must equal sqrtf(x) for every float in [0, 2^64) (GPU only: it
  corrects the hardware's square-root estimate).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from physica_amd import build

SUPPORT_RADII = [0.02, 0.04, 0.025, 1.0 / 30.0, 0.1, 0.0123]


def _library():
    if not build.MATH_CHECK.exists():
        build.build_math_check()
    lib = ctypes.CDLL(str(build.MATH_CHECK))
    lib.phx_test_div_host.restype = ctypes.c_uint64
    lib.phx_test_div_host.argtypes = [ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32]
    lib.phx_test_math_gpu.restype = ctypes.c_int
    lib.phx_test_math_gpu.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                      ctypes.POINTER(ctypes.c_ulonglong)]
    return lib


def _f32(x) -> float:
    return float(np.float32(x))


def _bits(x) -> int:
    return int(np.array([x], dtype=np.float32).view(np.uint32)[0])


def _divisors():
    out = []
    for h in SUPPORT_RADII:
        h32 = np.float32(h)
        out.append(_f32(h32))
        # the viscosity Laplacian's constant, pi h^6 in the same FP32 order
        out.append(_f32(np.float32(np.pi) * h32 * h32 * h32 * h32 * h32 * h32))
    rng = np.random.default_rng(20261202)
    out += [_f32(v) for v in rng.uniform(1e-3, 1.0, 8)]
    return out


@pytest.mark.parametrize("b", _divisors())
def test_div_rn_matches_ieee_division_over_two_binades(b):
    lib = _library()
    start = _bits(b) & 0xFF800000  # first float of b's binade
    for first in (start - (1 << 23), start - (2 << 23)):
        assert lib.phx_test_div_host(b, first, 1 << 23) == 0, (b, hex(first))


def test_div_rn_zero_and_the_divisor_itself():
    lib = _library()
    for b in _divisors():
        assert lib.phx_test_div_host(b, 0, 1) == 0  # x = +0
        assert lib.phx_test_div_host(b, _bits(b), 1) == 0  # x = b


def test_div_rn_check_catches_a_wrong_quotient():
    # the check itself: a one-ulp-off reciprocal must make plain x * inv_b
    # disagree with x / b somewhere in a binade, which div_rn's refinement fixes
    b = np.float32(0.02)
    x = (np.arange(1 << 16, dtype=np.uint32) + np.uint32(_bits(0.01))).view(np.float32)
    naive = x * (np.float32(1.0) / b)
    assert np.any(naive.view(np.uint32) != (x / b).view(np.uint32))


@pytest.mark.gpu
def test_sqrt_rn_equals_sqrtf_below_2_64():
    lib = _library()
    wrong = ctypes.c_ulonglong(0)
    last = _bits(2.0 ** 64)
    assert lib.phx_test_math_gpu(0, 0, last, 1.0, ctypes.byref(wrong)) == 0
    assert wrong.value == 0


def test_div_rn_domain_boundary():
    # every x in [2^-100, 2^-90) is exact; misses start below 2^-105, where
    # the remainders underflow
    lib = _library()
    b = _f32(0.02)
    low = _bits(2.0 ** -100)
    assert lib.phx_test_div_host(b, low, (_bits(2.0 ** -90) - low)) == 0
    assert lib.phx_test_div_host(b, _bits(2.0 ** -110), 1 << 23) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("h", [0.02, 0.04, 1.0 / 30.0])
def test_div_rn_equals_division_for_every_distance_below_h(h):
    lib = _library()
    wrong = ctypes.c_ulonglong(0)
    b = _f32(h)
    low = _bits(2.0 ** -100)
    assert lib.phx_test_math_gpu(1, low, _bits(b) - low, b, ctypes.byref(wrong)) == 0
    assert wrong.value == 0
    assert lib.phx_test_math_gpu(1, 0, 1, b, ctypes.byref(wrong)) == 0  # x = +0
    assert wrong.value == 0
