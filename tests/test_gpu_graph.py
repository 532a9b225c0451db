"""RK4 steps replayed from a captured HIP graph (wtg_set_graph) against the
same steps launched eagerly and against the oracle (needs a GPU).

Bar: bit-identical to the eager launches, FP64 and FP32-bulk, across
replays with eager steps between them and after a change of dt (which
captures the graph again); the FP64 graph path also equals the oracle bit for
bit.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import cases
from oracle.oracle import Oracle, geometry_of

pytestmark = pytest.mark.gpu


def _ctx(g, L, topo, m, t, u, precision, graph):
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.set_precision(precision)
    if precision == W.PRECISION_FP32_BULK:
        ctx.set_stage_shape(W.SHAPE_LAYER_PAIRS)
    ctx.set_graph(graph)
    return ctx


def _same(a, b, own, own_e):
    return (np.array_equal(a[0][:, own], b[0][:, own]) and np.array_equal(a[1][:, own], b[1][:, own])
            and np.array_equal(a[2][:, :, own_e], b[2][:, :, own_e]))


@pytest.mark.parametrize("precision", [W.PRECISION_FP64, W.PRECISION_FP32_BULK])
def test_graph_steps_equal_eager_steps(precision):
    print(f"graph vs eager, precision {precision}, J5 L4", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L = 4
    topo, m, t, u = cases.ocean(g, L, seed=7)
    own, own_e = g.layout.owned_cells(), g.layout.owned_edge_cells()
    eager = _ctx(g, L, topo, m, t, u, precision, False)
    graph = _ctx(g, L, topo, m, t, u, precision, True)
    # replays, eager steps in between, an odd step count, then a new dt
    for nsteps, dt, mode in ((3, 600.0, "graph"), (1, 600.0, "eager"), (4, 600.0, "graph"), (2, 450.0, "graph")):
        eager.rk4(0, dt, nsteps)
        if mode == "eager":
            graph.set_graph(False)
        graph.rk4(0, dt, nsteps)
        graph.set_graph(True)
        assert _same(eager.download_state(0), graph.download_state(0), own, own_e), (nsteps, dt, mode)
    if precision == W.PRECISION_FP64:
        o = Oracle(g.layout, geometry_of(g), topo)
        mm, tt, uu = m.copy(), t.copy(), u.copy()
        o.ghost_fill(mm, tt, uu)
        mm, tt, uu = o.rk4(mm, tt, uu, 600.0, 8)
        want = o.rk4(mm, tt, uu, 450.0, 2)
        assert _same(graph.download_state(0), want, own, own_e)
