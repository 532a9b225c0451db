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
