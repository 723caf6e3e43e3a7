"""The variation streams on the host (the device kernel's own functions): Xoshiro256++ and
its SplitMix64 seeding against the algorithms' published known answers, and the samplers'
documented constructions over those raw outputs."""
import ctypes as C

import numpy as np
import pytest

from rsr_amd import api

U64 = C.POINTER(C.c_uint64)


def _lib():
    l = api.lib()
    l.rsr_xoshiro256pp.argtypes = [U64, C.c_int, U64]
    l.rsr_xoshiro_seed.argtypes = [C.c_uint64, U64]
    l.rsr_vary_samples.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint32,
                                   C.POINTER(C.c_float)]
    return l


def raw(state, n):
    s = np.array(state, np.uint64)
    out = np.zeros(n, np.uint64)
    assert _lib().rsr_xoshiro256pp(s.ctypes.data_as(U64), n, out.ctypes.data_as(U64)) == 0
    return [int(x) for x in out]


def seeded(seed):
    s = np.zeros(4, np.uint64)
    assert _lib().rsr_xoshiro_seed(seed, s.ctypes.data_as(U64)) == 0
    return [int(x) for x in s]


def samples(seed, kind, n, a=0.0, b=1.0, bound=1):
    out = np.zeros(n, np.float32)
    assert _lib().rsr_vary_samples(seed, kind, n, a, b, bound, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return out


def test_xoshiro256pp_known_answer():
    """Blackman & Vigna's xoshiro256++ from state {1, 2, 3, 4} (the rand_xoshiro test vector)."""
    assert raw([1, 2, 3, 4], 10) == [41943041, 58720359, 3588806011781223, 3591011842654386,
                                     9228616714210784205, 9973669472204895162, 14011001112246962877,
                                     12406186145184390807, 15849039046786891736, 10450023813501588000]


def test_splitmix64_seeding_known_answer():
    """seed_from_u64(0) fills the state with SplitMix64's first four outputs from 0."""
    assert seeded(0) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


@pytest.mark.parametrize("seed", [0, 42, 2**63 + 5])
def test_f32_and_bool_from_upper_bits(seed):
    """f32 = (upper 32 bits >> 8) * 2^-24; bool = the top bit."""
    r = raw(seeded(seed), 64)
    f = samples(seed, 0, 64)
    assert np.array_equal(f, np.array([((x >> 32) >> 8) * 2.0**-24 for x in r], np.float32))
    bits = samples(seed, 1, 64)
    assert np.array_equal(bits, np.array([(x >> 63) & 1 for x in r], np.float32))


def test_below_in_range_and_uniform():
    for n in (1, 2, 3, 7, 25, 1000):
        v = samples(9, 2, 20000, bound=n)
        assert v.min() >= 0 and v.max() <= n - 1 and np.all(v == np.floor(v))
        if n <= 25:
            counts = np.bincount(v.astype(int), minlength=n)
            assert counts.min() > 0.8 * 20000 / n and counts.max() < 1.2 * 20000 / n, (n, counts)


@pytest.mark.parametrize("a,b", [(-5.0, 5.0), (0.9, 1.1), (-0.1, 0.1), (-0.5, 0.5)])
def test_range_bounds(a, b):
    v = samples(11, 3, 20000, a, b)
    assert np.all(v >= np.float32(a)) and np.all(v < np.float32(b))
    assert abs(float(v.mean()) - (a + b) / 2) < 0.02 * (b - a)
