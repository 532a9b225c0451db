"""The ghost fill split into local and would-be-remote launches (sharding
model, whole-lozenge shards) equals the full ghost fill bit for bit (needs a
GPU).  The slot filled is a trend, whose halo is zero before the fill."""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nshards", [1, 2, 5, 10])
def test_split_fill_equals_full_fill(nshards):
    print(f"split ghost fill, {nshards} shards, J5 L3", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    topo, m, t, u = cases.ocean(g, 3, seed=4)
    out = []
    for split in (False, True):
        ctx = W.Context(g, nlayers=3)
        ctx.set_topography(topo)
        ctx.upload_state(0, m, t, u)
        ctx.trend(0, 1)  # slot 1: trend on owned cells and poles, zero halo
        before = ctx.download_state(1)
        if split:
            ms_local, ms_remote = ctx.ghost_fill_split_timed(1, nshards, reps=3)
        else:
            ctx.ghost_fill(1)
        out.append(ctx.download_state(1))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    assert not np.array_equal(before[0], out[1][0])  # the fill changed the halo
    counts = g.shard_halo(nshards)
    assert (ms_remote == 0.0) == (counts[1] + counts[3] == 0)
    assert (ms_local == 0.0) == (counts[0] + counts[2] == 0)
