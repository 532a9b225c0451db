"""Sharding model of the ghost fill (whole-lozenge shards): the library's
counts of local and would-be-remote ghost entries against a count made here
from the grid's own ghost maps (node_src, edge_src), and the identities they
must satisfy (CPU only)."""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W


def _count(g, nshards):
    lay = g.layout
    S, n, nc = lay.S, lay.n, lay.ncells
    cells = np.arange(nc)
    lz, r = cells // (S * S), cells % (S * S)
    j, i = r // S - 2, r % S - 2
    halo = ~((i >= 0) & (i < n) & (j >= 0) & (j < n))
    shard = lz * nshards // 10
    src = g.array("node_src")
    nd = halo & (src >= 0) & (src != cells)
    node_local = int(np.sum(nd & (shard[np.maximum(src, 0)] == shard)))
    esrc = g.array("edge_src").reshape(nc, 3)
    slots = 3 * cells[:, None] + np.arange(3)[None, :]
    ed = halo[:, None] & (esrc >= 0) & (esrc != slots)
    eshard = shard[np.maximum(esrc, 0) // 3]
    edge_local = int(np.sum(ed & (eshard == shard[:, None])))
    return node_local, int(nd.sum()) - node_local, edge_local, int(ed.sum()) - edge_local


@pytest.mark.parametrize("level", [4, 5])
def test_shard_halo_counts(level):
    g = W.Grid(level=level, min_level=level, domain_level=2)
    total = None
    for ns in (1, 2, 3, 5, 10):
        got = g.shard_halo(ns)
        assert got == _count(g, ns)
        t = (got[0] + got[1], got[2] + got[3])
        total = total or t
        assert t == total  # every ghost entry is either local or remote
    assert g.shard_halo(1)[1] == 0 and g.shard_halo(1)[3] == 0
    assert g.shard_halo(10)[0] == 0 and g.shard_halo(10)[2] == 0  # one lozenge per shard: all remote
    with pytest.raises(W.WtgError):
        g.shard_halo(11)
