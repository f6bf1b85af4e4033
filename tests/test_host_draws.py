"""The library's random streams (host only, no GPU) against libstdc++ itself: the oracle's
orc_lmc_samples and orc_rademacher run std::mt19937 with std::normal_distribution<float> and
std::uniform_real_distribution<float> compiled by g++."""
import time

import numpy as np
import pytest

import oracle
from slemm_amd import gpu


@pytest.mark.parametrize("seed,m,n,nmc", [(0, 1, 1, 1), (0, 7, 5, 3), (3, 1001, 333, 4), (11, 40_000, 9_999, 3),
                                          (-54321, 100, 100, 2)])
def test_lmc_draws_equal_libstdcxx_normal_distribution(seed, m, n, nmc):
    u, e = gpu.lmc_draws(seed, m, n, nmc)
    ru, re = oracle.lmc_samples(seed, n, m, nmc)
    assert np.array_equal(u.view(np.uint32), ru.view(np.uint32))
    assert np.array_equal(e.view(np.uint32), re.view(np.uint32))


def test_lmc_draws_at_full_size_and_empty():
    """n = 100k, m = 500k, 3 samples (the reference's n_MC there): 1.8M draws, bit for bit; and
    zero-sized blocks."""
    t0 = time.perf_counter()
    u, e = gpu.lmc_draws(0, 500_000, 100_000, 3)
    t1 = time.perf_counter()
    ru, re = oracle.lmc_samples(0, 100_000, 500_000, 3)
    t2 = time.perf_counter()
    print(f"library {t1 - t0:.3f} s, libstdc++ {t2 - t1:.3f} s")
    assert np.array_equal(u.view(np.uint32), ru.view(np.uint32))
    assert np.array_equal(e.view(np.uint32), re.view(np.uint32))
    u0, e0 = gpu.lmc_draws(0, 0, 5, 2)
    assert u0.shape == (0, 2) and np.array_equal(e0, oracle.lmc_samples(0, 5, 0, 2)[1])


@pytest.mark.parametrize("seed,n,nv", [(0, 1, 1), (0, 5, 3), (7, 1000, 4), (0, 100_000, 7), (2, 623, 2), (2, 625, 1)])
def test_rademacher_probes_equal_libstdcxx(seed, n, nv):
    """Probe entries bit for bit (signs and the normalisation) across MT19937 block edges."""
    got = gpu.rademacher_probes(seed, n, nv)
    ref = oracle.rademacher(seed, n, nv)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_rademacher_sign_is_an_integer_compare():
    """The library takes a draw as -1 when its engine word u < 2^31 - 64, in place of
    generate_canonical<float, 24>(u) = float(u) / 2^32 < 0.5: float(u) rounds to nearest even, so
    from the midpoint 2^31 - 64 of the floats 2^31 - 128 and 2^31 on it reaches 2^31.  Checked on
    every word within 2^16 of 2^31 and on 2^22 random words."""
    near = np.arange(2**31 - 2**16, 2**31 + 2**16, dtype=np.uint64).astype(np.uint32)
    rnd = np.random.default_rng(1).integers(0, 2**32, size=2**22, dtype=np.uint64).astype(np.uint32)
    for u in (near, rnd):
        canon = u.astype(np.float32) * np.float32(2.0**-32)
        assert np.array_equal(canon < np.float32(0.5), u < np.uint32(2**31 - 64))
    assert (np.float32(np.uint32(2**31 - 65)) < 2**31) and not (np.float32(np.uint32(2**31 - 64)) < 2**31)


def test_rademacher_probes_in_parallel_chunks_equal_libstdcxx():
    """Streams of at least two chunks of 624 * 2^14 draws are produced by host threads from jumped
    engine states (21,000,030 draws: two chunks, the second to the stream's end, a last partial
    byte of sign bits); bit for bit the single libstdc++ stream."""
    got = gpu.rademacher_probes(0, 700_001, 30)
    ref = oracle.rademacher(0, 700_001, 30)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
