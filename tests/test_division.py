"""The DDA's division by a ray's direction component uses a reciprocal plus
one FMA correction (Markstein) instead of an IEEE division, and reciprocals
use the hardware reciprocal plus one FMA Newton step; these check both equal
IEEE division bit for bit."""
import ctypes as C
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def divcheck():
    path = os.path.join(HERE, "libdivcheck.so")
    if not os.path.exists(path):
        import __graft_entry__ as ge
        ge.build_oracle()
    lib = C.CDLL(path)
    lib.divcheck.restype = C.c_long
    lib.divcheck.argtypes = [C.c_long, C.c_uint64, C.c_int]
    return lib


def test_cpu_random_signed_scaled_pairs(divcheck):
    assert divcheck.divcheck(20_000_000, 12345, 0) == 0


def test_cpu_pairs_next_to_rounding_midpoints(divcheck):
    assert divcheck.divcheck(20_000_000, 777, 1) == 0


@pytest.mark.gpu
def test_gpu_all_numerators_for_spread_divisors(gpu):
    """Every numerator significand against 2^16 divisor significands spread over
    [1, 2) (including 1 and the all-ones significand), on the device."""
    ctx = gpu.Context((8, 8, 8))
    bad = 0
    # strided blocks of 64 consecutive divisors: 1024 blocks x 64 = 2^16 divisors
    for b in range(1024):
        bad += ctx.selftest_division(b * 8192, 64)
    bad += ctx.selftest_division((1 << 23) - 64, 64)
    assert bad == 0
    ctx.close()


@pytest.mark.gpu
def test_gpu_reciprocal_every_float(gpu):
    """The reciprocal path against IEEE 1/d for all 2^32 float bit patterns."""
    ctx = gpu.Context((8, 8, 8))
    assert ctx.selftest_reciprocal(0, 1 << 32) == 0
    ctx.close()
