"""Host side of the grid support in the product library (no GPU): ConstIndex,
is_intersect and the local-array offsets of a block-cyclic layout, checked
against the oracle's restatements and against brute-force ownership."""
import numpy as np
import pytest

from psmma_git_amd import api


def owner_lists(n, nb, npcol):
    """Brute force: global (1-based) indices owned by each process column."""
    own = [[] for _ in range(npcol)]
    for g in range(n):
        own[(g // nb) % npcol].append(g + 1)
    return own


@pytest.mark.parametrize("n,nb,npcol", [(1, 1, 1), (61, 8, 2), (61, 5, 3), (61, 16, 2),
                                        (61, 7, 4), (2048, 64, 2), (100, 100, 3), (5, 8, 4),
                                        (0, 4, 2)])
def test_constindex_matches_oracle_and_ownership(ora, n, nb, npcol):
    own = owner_lists(n, nb, npcol)
    for c in range(npcol):
        got = api.constindex(n, nb, c, npcol)
        assert got.tolist() == own[c]
        assert got.tolist() == ora.constindex(n, nb, c, npcol).tolist()
        assert len(got) == ora.numroc(n, nb, c, npcol)


def test_constindex_rejects_invalid():
    for args in [(-1, 4, 0, 2), (8, 0, 0, 2), (8, 4, 2, 2), (8, 4, -1, 2), (8, 4, 0, 0)]:
        with pytest.raises(ValueError):
            api.constindex(*args)


def test_is_intersect_reference_sense(ora):
    # TRUE iff disjoint (SRC/Auxil_2p5D.f90:772-787), on every pair of process columns
    n, nb, q = 61, 5, 3
    lists = [api.constindex(n, nb, c, q) for c in range(q)]
    for a in range(q):
        for b in range(q):
            want = a != b
            assert api.is_intersect(lists[a], lists[b]) == want
            assert ora.is_intersect(lists[a], lists[b]) == want
    assert api.is_intersect([1, 4, 9], [2, 3, 9, 10])  is False
    assert api.is_intersect([1, 4, 9], [2, 3, 8, 10]) is True
    assert api.is_intersect([], [1, 2]) is True


@pytest.mark.parametrize("m,n,p,q,mb,nb", [(61, 47, 2, 3, 8, 5), (64, 64, 2, 2, 32, 32),
                                           (10, 7, 3, 1, 4, 4), (5, 5, 4, 4, 2, 3)])
def test_local_offsets_tile_the_matrix(ora, m, n, p, q, mb, nb):
    off = api.local_offsets(m, n, p, q, mb, nb)
    assert off[0] == 0 and off[-1] == m * n
    for pr in range(p):
        for pc in range(q):
            r = pr * q + pc
            assert off[r + 1] - off[r] == ora.numroc(m, mb, pr, p) * ora.numroc(n, nb, pc, q)
