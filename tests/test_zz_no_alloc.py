"""cfpq_add_mul_noalloc (the add_mul that allocates no device memory) against
cfpq_add_mul, through the C ABI (needs a GPU).

Bar: bit-exact.  Two calls in a row, the second on a smaller matrix, so that
whatever the first call left in the shared scratch array must not reach the
second's result; then aliasing, every kernel path, and the size limit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(eng, oracle, rng, n, density):
    """The same random (c, a, b) uploaded twice: one set per entry point."""
    a, b, c = (rng.random((n, n)) < d for d in (density, density, density / 4))
    sets = []
    for _ in range(2):
        ms = []
        for d in (c, a, b):
            m = eng.matrix(n)
            m.upload(oracle.from_dense(d))
            ms.append(m)
        sets.append(ms)
    return sets


def _check(eng, oracle, rng, n, density):
    (c0, a0, b0), (c1, a1, b1) = _pair(eng, oracle, rng, n, density)
    want_changed = eng.add_mul(c0, a0, b0)
    got_changed = eng.add_mul_noalloc(c1, a1, b1)
    want, got = c0.download(), c1.download()
    assert np.array_equal(got, want), (n, density)
    assert got_changed == want_changed
    assert c1.fingerprint() == c0.fingerprint()
    for m in (c0, a0, b0, c1, a1, b1):
        m.free()


def test_noalloc_matches_add_mul_two_calls_smaller_second(pkg, oracle):
    print("[no_alloc] max size then a smaller one", flush=True)
    rng = np.random.default_rng(2024)
    with pkg.Engine(0) as eng:
        _check(eng, oracle, rng, pkg.NOALLOC_MAX_N, 0.002)   # the largest size the scratch array holds
        _check(eng, oracle, rng, 700, 0.01)                  # smaller: stale scratch must not leak in
        _check(eng, oracle, rng, 1, 1.0)
        _check(eng, oracle, rng, 257, 0.3)                   # dense tiles: the LDS / Four Russians paths


@pytest.mark.parametrize("kernel", ["gather", "lds", "m4r", "mfma"])
def test_noalloc_every_kernel(pkg, oracle, kernel):
    print(f"[no_alloc] kernel {kernel}", flush=True)
    rng = np.random.default_rng(7)
    with pkg.Engine(0, kernel=kernel) as eng:
        _check(eng, oracle, rng, 1000, 0.05)
        _check(eng, oracle, rng, 300, 0.2)


@pytest.mark.parametrize("alias", ["c=a", "c=b", "c=a=b"])
def test_noalloc_aliasing(pkg, oracle, alias):
    print(f"[no_alloc] aliasing {alias}", flush=True)
    n = 600
    rng = np.random.default_rng(11)
    c, o = rng.random((n, n)) < 0.01, rng.random((n, n)) < 0.01
    with pkg.Engine(0) as eng:
        res = []
        for fn in (eng.add_mul, eng.add_mul_noalloc):
            dc, do = eng.matrix(n), eng.matrix(n)
            dc.upload(oracle.from_dense(c))
            do.upload(oracle.from_dense(o))
            args = {"c=a": (dc, dc, do), "c=b": (dc, do, dc), "c=a=b": (dc, dc, dc)}[alias]
            ch = fn(*args)
            res.append((dc.download(), ch))
            dc.free()
            do.free()
        assert np.array_equal(res[0][0], res[1][0]) and res[0][1] == res[1][1]


def test_noalloc_rejects_large_and_mismatched(pkg):
    print("[no_alloc] argument errors", flush=True)
    with pkg.Engine(0) as eng:
        big = eng.matrix(pkg.NOALLOC_MAX_N + 1)
        with pytest.raises(pkg.CfpqError) as ei:
            eng.add_mul_noalloc(big, big, big)
        assert ei.value.code == pkg.E_INVALID and "limit" in str(ei.value)
        small = eng.matrix(10)
        with pytest.raises(pkg.CfpqError) as ei:
            eng.add_mul_noalloc(small, big, small)
        assert ei.value.code == pkg.E_INVALID
        # the context and the entry point stay usable after the errors
        x = eng.matrix(10)
        assert eng.add_mul_noalloc(x, small, small) is False
        for m in (big, small, x):
            m.free()
