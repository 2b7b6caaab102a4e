"""The find-by-id oracle against a linear scan and the reference's
documented example (tests/id_search.h:52-100)."""
import numpy as np

from oracle import idmap


def test_documentation_example():
    """N = 100 particles with ids 0..99 shuffled: find(2) is the particle with
    id 2, find(2N) is the end of the container."""
    n = 100
    ids = np.random.default_rng(0).permutation(n).astype(np.int64)
    k, v = idmap.build(ids)
    i2 = idmap.find(k, v, [2])[0]
    assert ids[i2] == 2
    assert idmap.find(k, v, [2 * n])[0] == n


def test_against_linear_scan():
    rng = np.random.default_rng(1)
    for n in (0, 1, 7, 1000):
        ids = rng.integers(0, max(n, 1) * 2, n).astype(np.int64)  # duplicates and gaps
        if n > 3:
            ids[0] = -5            # an id with the top bit set (a huge size_t)
            ids[1] = 1 << 40       # an id with high bits
            ids[2] = ids[3]        # a duplicate
        q = np.concatenate([ids, rng.integers(-10, 3 * max(n, 1), 200), [1 << 40, -5, 1 << 62]])
        k, v = idmap.build(ids)
        assert np.all(np.diff(k.astype(np.float64)) >= 0) or n < 2
        assert np.array_equal(idmap.find(k, v, q), idmap.find_brute(ids, q))
