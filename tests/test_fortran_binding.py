"""The Fortran ISO_C_BINDING module (wavetrisk_hydrostatic_amd/fortran/
wavetrisk_gpu.f90) calls the library as a WAVETRISK caller would.

The package build compiles the module and tests/fortran/binding_check.f90
with amdflang.  The program reads a level range, parameters and per-level
fields from a file, calls the C ABI through the module and writes what it
got back.
  grid mode (no GPU)         wtg_grid_info and wtg_grid_get("hex_inv") equal
                             what the ctypes binding reads
  rk4 mode (GPU)             wtg_rk4 on one level equals the ctypes path's
                             result bit for bit
  rk4d mode (GPU)            the same with wtg_set_diffusion (hyperdiffusion
                             of every order 2) set through the module's
                             wtg_diffusion type
  ml mode (GPU)              wtg_ml_rk4_split on a J4..J5 hierarchy with the
                             FMG free surface (wtg_mg_params from
                             wtg_mg_default_params, tolerances 1e-9) equals
                             the ctypes path bit for bit, iteration count
                             included
"""
import os
import subprocess

import numpy as np
import pytest

import wavetrisk_hydrostatic_amd as W
from wavetrisk_hydrostatic_amd import build as B
from wavetrisk_hydrostatic_amd import cases

EXE = os.path.join(B.FORTRAN_BIN, "binding_check")
needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="amdflang was not available to the build")


def _write(path, levels, min_level, domain_level, L, nsteps, par, fields):
    with open(path, "wb") as f:
        np.array([levels[0], levels[-1], min_level, domain_level, L, nsteps], dtype=np.int32).tofile(f)
        np.array(par, dtype=np.float64).tofile(f)
        for topo, m, t, u, e in fields:
            for a in (topo, m, t, u, e):
                np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def _run(mode, inp, out):
    r = subprocess.run([EXE, mode, inp, out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def _read(path, grids, L):
    raw = open(path, "rb").read()
    off, res = 0, []
    for g in grids:
        nc = g.layout.ncells
        parts = []
        for n in (L * nc, L * nc, 3 * L * nc, nc):
            parts.append(np.frombuffer(raw, dtype=np.float64, count=n, offset=off))
            off += 8 * n
        res.append((parts[0].reshape(L, nc), parts[1].reshape(L, nc), parts[2].reshape(L, 3, nc), parts[3]))
    it = int(np.frombuffer(raw, dtype=np.int32, count=1, offset=off)[0])
    return res, it


@needs_exe
def test_grid_through_fortran(tmp_path):
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write(inp, [4, 5], 4, 2, 1, 0, [0.0, 0.0, 0.0], [])
    _run("grid", inp, out)
    raw = open(out, "rb").read()
    off = 0
    for lev in (4, 5):
        g = W.Grid(level=lev, min_level=4, domain_level=2)
        n, S, halo = np.frombuffer(raw, dtype=np.int32, count=3, offset=off)
        nc, sp, npole = np.frombuffer(raw, dtype=np.int64, count=3, offset=off + 12)  # stream: no padding
        off += 36
        lay = g.layout
        assert (n, S, nc, sp, npole) == (lay.n, lay.S, lay.ncells, lay.south_pole, lay.north_pole)
        hexinv = np.frombuffer(raw, dtype=np.float64, count=nc, offset=off)
        off += 8 * nc
        assert np.array_equal(hexinv, g.array("hex_inv"), equal_nan=True)


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["rk4", "rk4d"])
def test_rk4_through_fortran_equals_ctypes(tmp_path, mode):
    print(f"Fortran binding: wtg_rk4 J5 L4 ({mode})", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L = 4
    topo, m, t, u = cases.ocean(g, L, seed=3)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write(inp, [5], 5, 2, L, 2, [600.0, 1.0, 1.0], [(topo, m, t, u, np.zeros(g.layout.ncells))])
    _run(mode, inp, out)
    res, _ = _read(out, [g], L)
    fm, ft, fu, _ = res[0]
    ctx = W.Context(g, nlayers=L, nslots=2)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    if mode == "rk4d":
        ctx.set_diffusion((2, 2, 2), (4e16, 2e16, 6e16, 4e16))
    ctx.rk4(0, 600.0, 2)
    cm, ct, cu = ctx.download_state(0)
    assert np.array_equal(fm, cm) and np.array_equal(ft, ct) and np.array_equal(fu, cu)


@needs_exe
@pytest.mark.gpu
def test_ml_split_through_fortran_equals_ctypes(tmp_path):
    print("Fortran binding: wtg_ml_rk4_split J4..J5 L2", flush=True)
    grids = [W.Grid(level=l, min_level=4, domain_level=2) for l in (4, 5)]
    L, dt, th1, th2 = 2, 600.0, 0.6, 0.8
    fields = []
    for g in grids:
        topo, m, t, u = cases.ocean(g, L, seed=5)
        fields.append((topo, m, t, u, m.sum(0) / 1028.0 + topo))
    # the coarse level restricted from the finest, as the reference initialises a hierarchy
    ctxs = []
    for g, fl in zip(grids, fields):
        c = W.Context(g, nlayers=L, nslots=2)
        c.set_topography(fl[0])
        c.fs_set_params(dt, th1, th2)
        ctxs.append(c)
    h = W.Hierarchy(ctxs)
    ml = W.MultiLevel(ctxs)
    try:
        ml.upload(0, *[[fl[k] for fl in fields] for k in (1, 2, 3)])
        ml.forward(0)
        ms, ts, us = ml.download(0)
        fields = [(fl[0], ms[i], ts[i], us[i], ms[i].sum(0) / 1028.0 + fl[0]) for i, fl in enumerate(fields)]
        inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
        _write(inp, [4, 5], 4, 2, L, 1, [dt, th1, th2], fields)
        _run("ml", inp, out)
        got, fit = _read(out, grids, L)
        for c, fl in zip(ctxs, fields):
            c.fs_upload(W.FS_ETA, fl[4])
        ml.upload(0, [fl[1] for fl in fields], [fl[2] for fl in fields], [fl[3] for fl in fields])
        it = ml.rk4_split(h, 0, 1, W.MgParams.defaults(coarse_tol=1e-9, fine_tol=1e-9))
        cm, ct, cu = ml.download(0)
        assert fit == it
        for i, c in enumerate(ctxs):
            fm, ft, fu, fe = got[i]
            assert np.array_equal(fm, cm[i]) and np.array_equal(ft, ct[i]) and np.array_equal(fu, cu[i])
            assert np.array_equal(fe, c.fs_download(W.FS_ETA))
    finally:
        ml.close()
        h.close()


@needs_exe
@pytest.mark.gpu
def test_vertical_diffusion_and_remap_through_fortran_equals_ctypes(tmp_path):
    """wtg_vdiff filled and changed through the Fortran derived type (a field
    out of place in the bind(C) type would change the result)"""
    print("Fortran binding: wtg_vertical_diffusion + wtg_remap (PPR) J5 L6", flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    L, NC = 6, g.layout.ncells
    topo, m, t, u = cases.ocean(g, L, seed=4)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write(inp, [5], 5, 2, L, 1, [900.0, 1.0, 1.0], [(topo, m, t, u, np.zeros(NC))])
    _run("vdr", inp, out)
    res, _ = _read(out, [g], L)
    fm, ft, fu, _ = res[0]
    ctx = W.Context(g, nlayers=L, nslots=2)
    ctx.set_topography(topo)
    ctx.upload_state(0, m, t, u)
    ctx.set_vdiff(tke_closure=1, Q_sr=20.0, bottom_friction=4e-4)
    ctx.vdiff_set_coefficients(np.full((L + 1, NC), 1.2e-5), np.full((L + 1, NC), 1.2e-4), np.full((L, NC), 2e-5))
    ctx.vdiff_set_forcing(np.full(NC, 1e-7), np.full(NC, 2e-7), np.full(NC, 0.1), np.full((3, NC), 1e-4))
    ctx.vertical_diffusion(0, 900.0, 1)
    k = np.arange(L + 1) / L
    b = 1.0 - k * k
    ctx.remap(0, W.REMAP_PPR, 1.0 - b, b)
    cm, ct, cu = ctx.download_state(0)
    ctx.close()
    assert np.array_equal(fm, cm) and np.array_equal(ft, ct) and np.array_equal(fu, cu)
    assert np.max(np.abs(cm - m)) > 0  # the remap moved the layers
