"""Multigrid with FP32 smoothing and FP64 residuals (Hierarchy.set_precision
(PRECISION_FP32_BULK)) on the GPU.

Each V-cycle is then a defect correction: the FP64 residual R = f - L u,
the correction equation L e = R smoothed, restricted and prolonged in FP32
fields, and u := u + e in FP64.  Bars:

* V-cycles on a J5..J7 draw reach the FP64 tolerance 1e-9 (the residual that
  ends the iteration is FP64) in at most one cycle more than the FP64
  cycles, and the two solutions agree within 1e-8 of the solution's maximum
  (both have residuals below 1e-9; the operator's conditioning bounds the
  difference);
* each FP32 cycle contracts the residual as the FP64 cycle does, to within
  a factor 1.05 in the first cycles (FP32 rounding of the correction is
  ~1e-7 relative), and the FP32 mode is in effect (residuals not identical);
* FMG on the reference's J4..J6 hierarchy (tests/golden/levels_J4.npz)
  reaches the tolerance on every level, with solutions within 1e-9 of the
  oracle's FP64 FMG solution;
* learned cycles (any network) with FP32 smoothing reach 1e-9;
* switching back to FP64 gives the FP64 cycles bit for bit.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from conftest import golden
from wavetrisk_hydrostatic_amd import learned as LC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob():
    p = LC.Problem(LC.hierarchy_grids(fine=7), seed=77)
    yield p
    p.close()


def _solve(prob, precision, method=W.SOLVE_VCYCLE, max_vcycle=30):
    prob.h.set_precision(precision)
    it, res, ms = prob.solve(method, LC.vcycle_params(1e-9, max_vcycle))
    return int(it[-1]), float(res[-1]), prob.top.fs_download(W.FS_ETA), ms


def test_fp32_vcycles_reach_fp64_tolerance(prob):
    print("FP32-smoothing V-cycles, J5..J7", flush=True)
    it64, r64, u64, ms64 = _solve(prob, W.PRECISION_FP64)
    it32, r32, u32, ms32 = _solve(prob, W.PRECISION_FP32_BULK)
    own = prob.grids[-1].layout.owned_cells()
    diff = np.max(np.abs(u32[own] - u64[own])) / np.max(np.abs(u64[own]))
    print(f"  FP64 {it64} cycles {r64:.2e} ({ms64:.1f} ms); FP32 smoothing {it32} cycles {r32:.2e} "
          f"({ms32:.1f} ms); solutions differ by {diff:.2e}", flush=True)
    assert r64 <= 1e-9 and r32 <= 1e-9
    assert it32 <= it64 + 1
    assert diff <= 1e-8
    # per-cycle contraction
    for n in (1, 2, 3):
        _, a, _, _ = _solve(prob, W.PRECISION_FP64, max_vcycle=n)
        _, b, _, _ = _solve(prob, W.PRECISION_FP32_BULK, max_vcycle=n)
        print(f"  after {n} cycles: FP64 {a:.6e}, FP32 smoothing {b:.6e}", flush=True)
        assert b <= 1.05 * a and a <= 1.05 * b
        if n == 1:
            assert a != b
    # back to FP64: the same cycles bit for bit
    it2, r2, u2, _ = _solve(prob, W.PRECISION_FP64)
    assert it2 == it64 and r2 == r64 and np.array_equal(u2, u64)
    with pytest.raises(W.WtgError, match="precision"):
        prob.h.set_precision(5)


def test_fp32_learned_cycles(prob):
    print("FP32-smoothing learned cycles, J5..J7 (untrained network)", flush=True)
    mlp = W.MLP(max_batch=prob.h.npatches, seed=9)
    try:
        prob.h.set_learned(mlp)
        it, r, _, _ = _solve(prob, W.PRECISION_FP32_BULK, method=W.SOLVE_LEARNED)
        print(f"  {it} cycles, residual {r:.2e}", flush=True)
        assert r <= 1e-9 and it < 30
    finally:
        prob.h.set_learned(None)
        prob.h.set_precision(W.PRECISION_FP64)
        mlp.close()


def test_fp32_fmg_against_oracle():
    print("FP32-smoothing FMG, reference J4..J6 hierarchy", flush=True)
    from test_gpu_levels import build
    f = golden("levels_J4.npz")
    grids, F, mg, ctxs = build(f)
    tol = float(f["tol"])
    for k, c in enumerate(ctxs):
        c.fs_upload(W.FS_ETA, np.zeros(grids[k].layout.ncells))
        c.fs_upload(W.FS_RHS, mg.levels[k].node_ghost(np.nan_to_num(F["f"][k])))
    h = W.Hierarchy(ctxs)
    try:
        h.set_precision(W.PRECISION_FP32_BULK)
        p = W.MgParams.defaults(coarse_tol=tol, fine_tol=tol, coarse_iter=400, max_vcycle=50)
        iters, res = h.solve(W.SOLVE_FMG, p)
    finally:
        h.close()
    print(f"  iterations {iters} (reference {list(f['fmg_iters'])}), residuals {res}", flush=True)
    assert np.all(res <= tol)
    zero = [np.zeros(g.layout.ncells) for g in grids]
    u_o, _, _ = mg.fmg(zero, [np.nan_to_num(x) for x in F["f"]])
    for k, g in enumerate(grids):
        o = g.layout.owned_cells()
        u = ctxs[k].fs_download(W.FS_ETA)
        scale = np.max(np.abs(u_o[k][o]))
        print(f"  level {k}: max difference to the oracle {np.max(np.abs(u[o] - u_o[k][o])) / scale:.2e}", flush=True)
        assert np.max(np.abs(u[o] - u_o[k][o])) <= 1e-9 * scale
