"""The allocation-free expectation entry point (qfvm_expect_paulis_noalloc) against
qfvm_expect_paulis, through the C-ABI: two calls in a row on one GPU, the second with a smaller
state and fewer terms (the fixed device scratch keeps the first call's contents), in complex128 and
complex64; results must be bit-identical (same kernels, same summation order)."""

import ctypes

import numpy as np
import pytest

from pyquafu_amd import _native
from tests.conftest import rand_state

pytestmark = pytest.mark.gpu


def expect_noalloc(dev, terms):
    c, p, o, l_ = _native.pauli_arrays(terms)
    out = np.full(len(o), np.nan)
    rc = _native.lib().qfvm_expect_paulis_noalloc(dev.handle, c.ctypes.data, p.ctypes.data, o.ctypes.data,
                                                  l_.ctypes.data, len(o), out.ctypes.data)
    _native.check(rc)
    return out


def random_terms(n, count, rng):
    terms = [("ZZ", [i, (i + 1) % n]) for i in range(n)] + [("X", [i]) for i in range(n)]
    terms.append(("X" * n, list(range(n))))  # wider than one tile: the fallback kernel
    while len(terms) < count:
        k = int(rng.integers(1, n + 1))
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        terms.append(("".join(rng.choice(list("IXYZ"), k)), qs))
    return terms


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_noalloc_matches_expect_paulis(dtype):
    rng = np.random.Generator(np.random.PCG64(4242))
    for n, count in ((20, 256), (11, 37)):
        dev = _native.DeviceState(n, dtype)
        try:
            dev.load(rand_state(n, rng))
            terms = random_terms(n, count, rng)
            got = expect_noalloc(dev, terms)  # first: before any other call that could allocate scratch
            want = dev.expect(terms)
            print(f"[noalloc {dtype}] n={n} terms={len(terms)} max |diff| {float(np.abs(got - want).max()):.1e}",
                  flush=True)
            assert np.array_equal(got, want)
        finally:
            dev.close()


def test_noalloc_rejects_too_many_terms():
    dev = _native.DeviceState(4, "c128")
    try:
        terms = [("Z", [k % 4]) for k in range(257)]
        with pytest.raises(_native.EngineError) as e:
            expect_noalloc(dev, terms)
        assert e.value.code == _native.ERR_VALUE and "at most 256 terms" in str(e.value)
    finally:
        dev.close()
