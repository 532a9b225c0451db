"""Grid generator (product host code) against the reference's own grid.

Fixtures grid_J4/J5 hold the reference's geometry on every cell the TRiSK
stencils read and its ghost maps; ghosts_J6/J7 hold the ghost maps of the
reference's param_J6 (domain level 3) and param_J7 (domain level 2) builds
(tests/golden/make_golden.py).
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from conftest import golden
from helpers import grid_of_fixture

# Cells where the reference stores values no stencil reads (its ghost
# geometry there is never computed or comes from uninitialised nodes):
# (hemisphere, i, j, array, component), i/j relative to n when given as "n+k".
UNUSED = {
    ("pedlen", -1, -1), ("len", "n+0", -1), ("len", "n+1", -1), ("len", -1, "n+0"), ("len", -1, "n+1"),
    ("area", "n+1", "n+1"),
}


def _key(n, i, j):
    f = lambda v: v if v < n - 2 else f"n{v - n:+d}".replace("+-", "-")
    return f(i), f(j)


@pytest.mark.parametrize("J", [4, 5, 6, 7])
def test_ghost_maps_match_reference(J):
    f = golden(f"grid_J{J}.npz" if J <= 5 else f"ghosts_J{J}.npz")
    g = grid_of_fixture(f)
    nc = g.layout.ncells
    ref_n, ref_e, ref_s = f["node_src"], f["edge_src"], f["edge_sign"]
    mine_n, mine_e, mine_s = g.array("node_src"), g.array("edge_src"), g.array("edge_sign")
    det = ref_n != -2
    assert det.sum() > 10 * 4 * g.layout.n  # every halo strip is covered
    assert np.array_equal(mine_n[det], ref_n[det])
    det = ref_e != -2
    assert det.sum() > 3 * 10 * 4 * g.layout.n
    assert np.array_equal(mine_e[det], ref_e[det])
    filled = det & (ref_e >= 0)
    assert np.array_equal(mine_s[filled], ref_s[filled])
    # the frozen (never filled) slots are exactly the reference's
    assert np.array_equal(mine_e[det] == -1, ref_e[det] == -1)


@pytest.mark.parametrize("J", [4, 5])
def test_geometry_matches_reference(J):
    f = golden(f"grid_J{J}.npz")
    g = grid_of_fixture(f)
    lay = g.layout
    lz, ii, jj = lay.coords()
    n = lay.n
    for name in ("len", "pedlen", "area", "hex_inv", "triarea", "coriolis"):
        ref = f["geo_" + name]
        mine = g.array(name)
        mine = mine.reshape(ref.shape)
        scale = np.nanmax(np.abs(ref), axis=1, keepdims=True)
        rel = np.abs(mine - ref) / scale
        bad = np.argwhere(np.isfinite(ref) & (rel > 1e-12))
        bad = [(comp, c) for comp, c in bad if (name,) + _key(n, ii[c], jj[c]) not in UNUSED]
        assert not bad, f"{name}: {len(bad)} cells differ, e.g. " + str(
            [(int(lz[c]), int(ii[c]), int(jj[c]), int(comp), ref[comp, c], mine[comp, c]) for comp, c in bad[:5]])


def test_grid_counts_and_closed_forms():
    """10*4^J+2 nodes; hexagon areas tile the sphere; Coriolis = 2 Omega z/R * area."""
    g = W.Grid(level=5, min_level=5, domain_level=2)
    lay = g.layout
    own = lay.owned_cells()
    assert len(own) == 10 * 4 ** 5 + 2
    area = 1.0 / g.array("hex_inv")[own]
    assert abs(area.sum() / (4 * np.pi * g.radius ** 2) - 1) < 1e-12
    parts = g.array("area")[:, own]
    assert np.allclose(parts.sum(axis=0) * g.array("hex_inv")[own], 1.0, rtol=1e-13)
    tri = g.array("triarea")[:, lay.owned_cells(with_poles=False)]
    assert abs(tri.sum() / (4 * np.pi * g.radius ** 2) - 1) < 1e-12
