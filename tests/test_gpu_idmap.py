"""GPU find-by-id map against the oracle (oracle/idmap.py), through the C
ABI.  Bar: sorted keys, values and every lookup bit-exact."""
import numpy as np
import pytest
import torch

import aboria_amd as ab
from oracle import idmap

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 100, 5000, 1_000_003])
def test_id_map_matches_oracle(ctx, n):
    print("id map n=%d" % n, flush=True)
    rng = np.random.default_rng(n + 3)
    ids = rng.integers(0, max(n, 1) * 2, n).astype(np.int64)
    if n > 3:
        ids[0], ids[1], ids[2] = -5, 1 << 40, ids[3]
    m = ab.IdMap(ctx, torch.tensor(ids, device="cuda"))
    k, v = idmap.build(ids)
    assert np.array_equal(m.keys().cpu().numpy().view(np.uint64), k)
    assert np.array_equal(m.values().cpu().numpy(), v)
    q = np.concatenate([ids, rng.integers(-10, 3 * max(n, 1), 5000), [1 << 40, -5, 1 << 62]])
    got = m.find(torch.tensor(q, device="cuda")).cpu().numpy()
    assert np.array_equal(got, idmap.find(k, v, q))
    m.close()


def test_id_search_after_reorder(ctx):
    """Particles.init_id_search + init_neighbour_search (the reference's
    helper_d_random, tests/id_search.h:193-300): after the reorder by bucket
    and the removal of dead particles, find(id) is the container index of
    the particle with that id (the brute-force scan), and the end for ids
    that are gone or never existed."""
    n = 20_000
    rng = np.random.default_rng(9)
    pos = torch.tensor(rng.uniform(-1, 1, (n, 3)), device="cuda")
    parts = ab.Particles(ctx, pos)
    parts.init_id_search()
    parts.alive[rng.choice(n, 300, replace=False)] = 0
    parts.init_neighbour_search([-1] * 3, [1] * 3, [False] * 3, 10.0)
    na = parts.size()
    assert na == n - 300
    ids = parts.id.cpu().numpy()
    q = rng.integers(0, n + 50, 4000)
    got = parts.find(torch.tensor(q, device="cuda")).cpu().numpy()
    assert np.array_equal(got, idmap.find_brute(ids, q))
    assert np.array_equal(parts.get_query().find(torch.tensor(q, device="cuda")).cpu().numpy(),
                          got)
    hit = got < na
    assert np.array_equal(ids[got[hit]], q[hit])
    # the documentation example: ids shuffled, find(2) has id 2, find(2N) is the end
    perm = rng.permutation(100)
    p2 = ab.Particles(ctx, torch.tensor(rng.random((100, 2)), device="cuda"))
    p2.id = torch.tensor(perm, dtype=torch.int64, device="cuda")
    p2.init_id_search()
    i2 = int(p2.find([2])[0])
    assert int(p2.id[i2]) == 2
    assert int(p2.find([200])[0]) == 100
