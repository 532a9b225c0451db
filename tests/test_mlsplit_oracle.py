"""The mode-split step on a full hierarchy against the reference.

Fixture tests/golden/mlsplit_J4.npz comes from the reference's RK4_split
(time_integr.f90:360-382, 444-478) with the FMG free-surface solver on the
hierarchy J4..J5 of its uniform refinement with zero thresholds
(oracle/ref_harness/dump_mlstep.f90): one layer, dt 600 s, theta1 0.6,
theta2 0.8, FMG to 1e-9, two steps; state and free surface on both levels
before and after each step.

Bars (relative to the field's max-norm over the level's owned cells), each
step started from the reference's own state before it:
  mass, velocity       within 1e-13 for |latitude| < 60 degrees
  free surface         within 1e-12 for |latitude| < 60 degrees
The reference keeps one pole copy per lozenge and lets them drift apart in
its transforms; this layout has one pole master per pole.  The elliptic
solve spreads the poles' difference over the sphere at the level of its
tolerance, so the caps are left out and the free surface's bar is wider.

Planted errors (in the oracle's parameters, on the CPU): theta2 = 1 (no old
flux divergence) and an undamped theta1 move the step far outside the bars.
"""
import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from conftest import golden
from oracle.oracle import MgOracle, MlOracle
from test_ml_oracle import outside_caps, owned

VARS = ("mass", "temp", "velo")


def load():
    f = golden("mlsplit_J4.npz")
    ls, le = int(f["level_start"]), int(f["level_end"])
    levels = list(range(ls, le + 1))
    grids = [W.Grid(level=l, min_level=int(f["min_level"]), domain_level=int(f["domain_level"]),
                    radius=float(f["radius"]), omega=float(f["omega"])) for l in levels]
    topos = [np.nan_to_num(f[f"topo_{l}"]) for l in levels]
    return f, levels, grids, topos


def mg_of(f, grids, topos, **over):
    kw = dict(grav=float(f["grav"]), rho0=float(f["ref_density"]), dt=float(f["dt"]), theta1=float(f["theta1"]),
              theta2=float(f["theta2"]), coarse_tol=float(f["tol"]), fine_tol=float(f["tol"]), coarse_iter=400,
              max_vcycle=50)
    kw.update(over)
    return MgOracle(grids, topos, **kw)


def state_before(f, levels, step):
    """the reference's state and free surface before `step` (1-based): its
    NaN slots (halo copies it never writes) taken from the initial state"""
    s0 = [[np.nan_to_num(f[f"s0_{v}_{l}"]) for l in levels] for v in VARS]
    eta = [np.nan_to_num(f[f"e{step - 1:02d}_eta_{l}"]).reshape(-1) for l in levels]
    if step == 1:
        return s0, eta
    st = [[f[f"s{step - 1:02d}_{v}_{l}"].copy() for l in levels] for v in VARS]
    for v in range(3):
        for i in range(len(levels)):
            nan = np.isnan(st[v][i])
            st[v][i][nan] = s0[v][i][nan]
    return st, eta


def errors(f, levels, grids, step, m, u, eta, cut=60.0):
    out = []
    for i, (l, g) in enumerate(zip(levels, grids)):
        oc, oe = outside_caps(g, cut)
        occ, oee = owned(g)
        rm, ru = f[f"s{step:02d}_mass_{l}"], f[f"s{step:02d}_velo_{l}"]
        re = f[f"e{step:02d}_eta_{l}"].reshape(-1)
        out.append((np.abs(m[i][:, oc] - rm[:, oc]).max() / np.abs(rm[:, occ]).max(),
                    np.abs(u[i][:, :, oe] - ru[:, :, oe]).max() / np.abs(ru[:, :, oee]).max(),
                    np.abs(eta[i][oc] - re[oc]).max() / np.abs(re[occ]).max()))
    return out


@pytest.fixture(scope="module")
def fx():
    f, levels, grids, topos = load()
    o = MlOracle(grids, topos, grav=float(f["grav"]), rho0=float(f["ref_density"]))
    return f, levels, grids, topos, o


@pytest.mark.parametrize("step", [1, 2])
def test_split_step_vs_reference(fx, step):
    f, levels, grids, topos, o = fx
    st, eta = state_before(f, levels, step)
    o.ghost_fill(*st)
    m, t, u, e, it = o.rk4_split(mg_of(f, grids, topos), *st, eta)
    assert (it >= 1).all() and (it <= 50).all()
    for l, (em, eu, ee) in zip(levels, errors(f, levels, grids, step, m, u, e)):
        assert em < 1e-13 and eu < 1e-13 and ee < 1e-12, (l, em, eu, ee)


def test_planted_parameter_errors_are_caught(fx):
    f, levels, grids, topos, o = fx
    st, eta = state_before(f, levels, 1)
    o.ghost_fill(*st)
    for over in (dict(theta2=1.0), dict(theta1=1.0)):
        m, t, u, e, it = o.rk4_split(mg_of(f, grids, topos, **over), *st, eta)
        worst = max(max(x) for x in errors(f, levels, grids, 1, m, u, e))
        assert worst > 1e-9, over
