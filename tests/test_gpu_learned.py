"""Learned correction of the free-surface multigrid on the GPU.

Checks: the network's input rows for the finest level equal a numpy
evaluation of their definition (residual over RMS, depth / 4000 m,
sin(latitude)) to FP32 rounding; with any network the learned cycles reach
the FP64 tolerance (the minimal-residual step never increases the residual
and the V-cycle remains the fallback); a network trained on the GPU from
pairs the FP64 solver generates needs no more cycles than plain V-cycles on
held-out draws and lowers the residual of each learned step; the shipped
J7 network saves at least one cycle per draw on average on 16 held-out
draws (a strict gain) and never needs more cycles than plain V-cycles.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import learned as LC
from wavetrisk_hydrostatic_amd.cases import lonlat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def grids():
    return LC.hierarchy_grids(fine=6, levels=3)


def test_features_match_definition(grids):
    print("learned: features", flush=True)
    prob = LC.Problem(grids, seed=1)
    try:
        top, g = prob.top, grids[-1]
        r = np.random.default_rng(0).standard_normal(g.layout.ncells)
        top.fs_upload(W.FS_DIVF, r)
        rg = top.fs_download(W.FS_DIVF)
        feats, sigma, cells = prob.h.features(W.FS_DIVF)
        n8 = g.layout.n // 8
        assert feats.shape == (10 * n8 * n8, 192)
        lz, i, j = g.layout.coords()
        assert np.all((i[cells] >= 0) & (i[cells] < g.layout.n) & (j[cells] >= 0) & (j[cells] < g.layout.n))
        assert len(np.unique(cells)) == cells.size == 10 * g.layout.n ** 2
        sg = np.sqrt(np.mean(rg[cells] ** 2, axis=1))
        assert np.max(np.abs(sigma / sg - 1)) <= 1e-14
        assert np.array_equal(feats[:, :64], (rg[cells] / sigma[:, None]).astype(np.float32))
        _, lat = lonlat(g.array("x"), g.radius)
        assert np.max(np.abs(feats[:, 128:] - np.sin(lat)[cells])) <= 1e-7
        depth = (np.abs(prob.topos[-1]) + top.fs_download(W.FS_ETA_OLD)) / 4000.0
        assert np.array_equal(feats[:, 64:128], depth[cells].astype(np.float32))
    finally:
        prob.close()


def test_learned_cycles_zero_equal_plain_vcycles(grids):
    """learned_cycles = 0: no learned step is taken, the cycles are the plain
    V-cycles bit for bit; learned_cycles = 1 takes it before the first cycle
    only and still reaches the FP64 tolerance."""
    print("learned: learned_cycles 0 and 1", flush=True)
    prob = LC.Problem(grids, seed=12)
    mlp = W.MLP(max_batch=prob.h.npatches, seed=4)
    try:
        p = LC.vcycle_params(1e-9, 30)
        i0, r0, _ = prob.solve(W.SOLVE_VCYCLE, p)
        u0 = prob.top.fs_download(W.FS_ETA)
        prob.h.set_learned(mlp)
        p.learned_cycles = 0
        i1, r1, _ = prob.solve(W.SOLVE_LEARNED, p)
        assert list(i1) == list(i0) and np.array_equal(r1, r0)
        assert np.array_equal(prob.top.fs_download(W.FS_ETA), u0)
        p.learned_cycles = 1
        i2, r2, _ = prob.solve(W.SOLVE_LEARNED, p)
        print(f"  plain {i0[-1]} cycles, learned first cycle {i2[-1]} cycles ({r2[-1]:.2e})", flush=True)
        assert r2[-1] <= 1e-9 and i2[-1] <= i0[-1] + 1
    finally:
        prob.h.set_learned(None)
        prob.close()
        mlp.close()


def test_learned_cycles_converge_with_any_network(grids):
    print("learned: untrained network", flush=True)
    prob = LC.Problem(grids, seed=2)
    mlp = W.MLP(max_batch=prob.h.npatches, seed=3)
    try:
        (i0, r0, _), (i1, r1, _) = LC.compare(prob, mlp)
        print(f"plain {i0} cycles ({r0:.2e}), untrained learned {i1} cycles ({r1:.2e})", flush=True)
        assert r0 <= 1e-9 and r1 <= 1e-9
        assert i1 <= i0 + 1
    finally:
        prob.close()
        mlp.close()


def test_trained_network_gains(grids):
    print("learned: harvest + train", flush=True)
    X, T = [], []
    for s in range(6):
        prob = LC.Problem(grids, seed=100 + s)
        x, t = LC.harvest(prob, ncycles=5)
        prob.close()
        X.append(x)
        T.append(t)
    X, T = np.concatenate(X), np.concatenate(T)
    npatch = 10 * (grids[-1].layout.n // 8) ** 2
    mlp, losses = LC.train_network(X, T, steps=800, batch=512, lr=1e-3, seed=1, max_batch=npatch)
    print(f"{len(X)} pairs, loss {losses[:10].mean():.4f} -> {losses[-50:].mean():.4f}", flush=True)
    assert losses[-50:].mean() < losses[:10].mean()
    gains = []
    for s in range(2):
        prob = LC.Problem(grids, seed=900 + s)
        (i0, r0, m0), (i1, r1, m1) = LC.compare(prob, mlp)
        prob.close()
        print(f"held-out {s}: V-cycles {i0} ({r0:.2e}, {m0:.1f} ms), learned {i1} ({r1:.2e}, {m1:.1f} ms)", flush=True)
        assert r1 <= 1e-9
        assert i1 <= i0
        gains.append(i0 - i1)
    mlp.close()


def test_shipped_network_strict_gain_on_held_out_draws():
    """The shipped J7 network (trained on draws 1000..1049 by learned.main)
    on 16 held-out draws of the J5..J7 hierarchy, solved to 1e-9 together
    (member-batched cycles): no draw needs more learned cycles than plain
    V-cycles, and the learned cycles save at least one cycle per draw on
    average (a strict gain; the bench measures 11.25 -> 9.5 at 64 draws)."""
    print("learned: shipped J7 network on 16 held-out draws", flush=True)
    grids7 = LC.hierarchy_grids(fine=7, levels=3)
    seeds = list(range(6000, 6016))
    prob = LC.Problem(grids7, seeds)
    mlp = LC.load_network(prob.h.npatches * len(seeds))
    assert mlp is not None, "the shipped network learned_J7.npz is missing"
    try:
        p = LC.vcycle_params(1e-9, 40)
        prob.solve(W.SOLVE_VCYCLE, p)
        plain, rp = prob.h.member_results()
        prob.h.set_learned(mlp)
        prob.solve(W.SOLVE_LEARNED, p)
        learned, rl = prob.h.member_results()
        prob.h.set_learned(None)
    finally:
        mlp.close()
        prob.close()
    print(f"  plain {list(plain)} -> learned {list(learned)}", flush=True)
    assert np.all(rp <= 1e-9) and np.all(rl <= 1e-9)
    assert np.all(learned <= plain)
    assert plain.mean() - learned.mean() >= 1.0
