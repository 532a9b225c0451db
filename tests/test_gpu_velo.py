"""HIP velocity restriction (Restrict_velo, wavelet.f90:1154-1174) against
the oracle (needs a GPU).

Bars: every owned coarse edge equals the oracle's restriction bit for bit
(the same mean of the same two fine edges, FP64, no contraction), all
layers; the restriction of a node field's discrete gradient is the coarse
gradient to 1e-12 (the closed form of tests/test_velo_restrict_oracle.py);
levels that are not consecutive are refused."""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from oracle.velo_oracle import restrict_velo
from test_velo_restrict_oracle import grad

pytestmark = pytest.mark.gpu


def test_restrict_velo_equals_oracle():
    print("velocity restriction J5 -> J4, 3 layers", flush=True)
    gf = W.Grid(level=5, min_level=4, domain_level=1)
    gc = W.Grid(level=4, min_level=4, domain_level=1)
    lf, lc = gf.layout, gc.layout
    L = 3
    rng = np.random.default_rng(3)
    fine, coarse = W.Context(gf, nlayers=L), W.Context(gc, nlayers=L)
    for c, lay in ((fine, lf), (coarse, lc)):
        c.set_topography(np.full(lay.ncells, -1000.0))
    uf = rng.standard_normal((L, 3, lf.ncells))
    fine.upload_state(0, np.full((L, lf.ncells), 1000.0), np.zeros((L, lf.ncells)), uf)
    coarse.upload_state(0, np.full((L, lc.ncells), 1000.0), np.zeros((L, lc.ncells)), np.zeros((L, 3, lc.ncells)))
    uf_filled = fine.download_state(0)[2]
    fine.restrict_velo(0, coarse, 0)
    got = coarse.download_state(0)[2]
    want = restrict_velo(lf, lc, uf_filled)
    own = lc.owned_cells(with_poles=False)
    assert np.array_equal(got[:, :, own], want[:, :, own])
    # gradient closed form through the device path
    xf, xc = gf.array("x").reshape(3, -1), gc.array("x").reshape(3, -1)
    phi_f, phi_c = np.cos(xf[1] / 6.371e6 * 3.0), np.cos(xc[1] / 6.371e6 * 3.0)
    g_f = np.repeat(grad(lf, gf.array("len").reshape(3, -1), phi_f)[None], L, axis=0)
    fine.upload_state(0, np.full((L, lf.ncells), 1000.0), np.zeros((L, lf.ncells)), g_f)
    fine.restrict_velo(0, coarse, 0)
    gc_want = grad(lc, gc.array("len").reshape(3, -1), phi_c)
    got = coarse.download_state(0)[2]
    err = np.max(np.abs(got[:, :, own] - gc_want[None][:, :, own])) / np.max(np.abs(gc_want[:, own]))
    print(f"  gradient closed form: {err:.2e}", flush=True)
    assert err <= 1e-12
    with pytest.raises(W.WtgError, match="coarser"):
        coarse.restrict_velo(0, fine, 0)
    fine.close()
    coarse.close()


def test_restrict_ss_samples_coincident_nodes():
    """Sub-sampling restriction (Restrict_ss, wavelet.f90:823-837): coarse
    nodes coincide with the fine level's even nodes, so restricting the fine
    node coordinates gives the coarse coordinates (poles included), and every
    coarse master and halo copy equals the fine value at its child."""
    print("sub-sampling restriction J5 -> J4", flush=True)
    grids = [W.Grid(level=l, min_level=4, domain_level=1) for l in (4, 5)]
    ctxs = []
    for g in grids:
        c = W.Context(g, nlayers=1)
        c.set_topography(np.full(g.layout.ncells, -1000.0))
        c.fs_set_params(100.0)
        ctxs.append(c)
    h = W.Hierarchy(ctxs)
    try:
        xf, xc = grids[1].array("x").reshape(3, -1), grids[0].array("x").reshape(3, -1)
        lc = grids[0].layout
        masters = lc.owned_cells()  # owned nodes and the two poles
        for d in range(3):
            ctxs[1].fs_upload(W.FS_DIVF, xf[d])
            h.restrict_ss(1, W.FS_DIVF, W.FS_DIVF_OLD)
            got = ctxs[0].fs_download(W.FS_DIVF_OLD)
            assert np.max(np.abs(got[masters] - xc[d][masters])) <= 1e-15 * 6.371e6
        # halo copies equal their masters
        src = grids[0].array("node_src")
        halo = np.nonzero((src >= 0) & (src != np.arange(len(src))))[0]
        assert np.array_equal(got[halo], got[src[halo]])
    finally:
        h.close()
        for c in ctxs:
            c.close()
