"""Find-by-id map restated in numpy.  TEST INFRASTRUCTURE ONLY: imported by
tests/ as the checker, never by the product package.

build() follows the id-map maintenance of neighbour_search_base::
update_positions (src/NeighbourSearchBase.h:440-486): keys = the ids of
the particles in container order, values = their container indices, then a
stable sort_by_key on the ids (size_t, i.e. unsigned 64-bit).  find()
follows CellListOrderedQuery::find (src/CellListOrdered.h:379-388): the
lower bound of the id among the keys; the value there if the key equals the
id, else n (the end of the container).
"""
import numpy as np


def build(ids):
    """(keys uint64, values int32) sorted by id, equal ids in index order."""
    u = np.ascontiguousarray(ids, np.int64).view(np.uint64)
    order = np.argsort(u, kind="stable")
    return u[order], order.astype(np.int32)


def find(keys, values, ids):
    """container index of each id (the lowest one), or len(keys)"""
    n = keys.shape[0]
    q = np.ascontiguousarray(ids, np.int64).view(np.uint64)
    pos = np.searchsorted(keys, q, side="left")
    at = np.minimum(pos, max(n - 1, 0))
    hit = (pos < n) & (keys[at] == q) if n else np.zeros(q.shape, bool)
    return np.where(hit, values[at] if n else 0, n).astype(np.int32)


def find_brute(ids, query):
    """linear scan: the lowest container index whose id equals each query
    (tests/id_search.h brute_force_check, first match), or len(ids)"""
    ids = np.asarray(ids, np.int64)
    out = np.full(len(query), len(ids), np.int32)
    for k, q in enumerate(np.asarray(query, np.int64)):
        hits = np.nonzero(ids == q)[0]
        if hits.size:
            out[k] = hits[0]
    return out
