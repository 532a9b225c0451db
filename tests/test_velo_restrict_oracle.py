"""The oracle's velocity restriction (Restrict_velo, wavelet.f90:1154-1174)
pinned by a closed form: fine edges bisect coarse edges exactly, so the
restriction of the discrete gradient of a node field (the edge differences
of grad_eta, ops.f90:217-220) is the coarse level's discrete gradient of the
same field at the coincident nodes, to rounding.  A wrong child edge, a
swapped neighbour or a flipped orientation moves the error to O(1)."""
import numpy as np

import wavetrisk_hydrostatic_amd as W
from oracle.velo_oracle import restrict_velo


def grad(layout, lens, phi):
    """edge differences over edge lengths, grad_eta's orientation (RT: E - c,
    DG: c - NE, UP: N - c) on owned nodes"""
    S, NC = layout.S, layout.ncells
    own = layout.owned_cells(with_poles=False)
    g = np.zeros((3, NC))
    g[0, own] = (phi[own + 1] - phi[own]) / lens[0, own]
    g[1, own] = (phi[own] - phi[own + S + 1]) / lens[1, own]
    g[2, own] = (phi[own + S] - phi[own]) / lens[2, own]
    return g


def test_restriction_of_a_gradient_is_the_coarse_gradient():
    gf = W.Grid(level=5, min_level=4, domain_level=1)
    gc = W.Grid(level=4, min_level=4, domain_level=1)
    lf, lc = gf.layout, gc.layout
    xf, xc = gf.array("x").reshape(3, -1), gc.array("x").reshape(3, -1)
    phi_f = np.sin(xf[0] / 6.371e6 * 2.0) + xf[2] / 6.371e6
    phi_c = np.sin(xc[0] / 6.371e6 * 2.0) + xc[2] / 6.371e6
    uf = grad(lf, gf.array("len").reshape(3, -1), phi_f)[None]
    want = grad(lc, gc.array("len").reshape(3, -1), phi_c)
    got = restrict_velo(lf, lc, uf)[0]
    own = lc.owned_cells(with_poles=False)
    scale = np.max(np.abs(want[:, own]))
    err = np.max(np.abs(got[:, own] - want[:, own])) / scale
    print(f"restricted gradient vs coarse gradient: {err:.2e}")
    assert err <= 1e-12
    # a planted mistake in the restatement (DG children swapped for N, NE) fails
    Sf = lf.S
    bad = restrict_velo(lf, lc, uf)
    lz, j, i = np.meshgrid(np.arange(10), np.arange(lc.n), np.arange(lc.n), indexing="ij")
    p = (lz * lc.S * lc.S + (j + 2) * lc.S + (i + 2)).ravel()
    c = (lz * Sf * Sf + (2 * j + 2) * Sf + (2 * i + 2)).ravel()
    bad[0, 1, p] = 0.5 * (uf[0, 1, c + Sf] + uf[0, 1, c])
    assert np.max(np.abs(bad[0][:, own] - want[:, own])) / scale > 1e-2
