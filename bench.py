"""Benchmark of the TRiSK hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--level 8] [--layers 30]

Workload (default): a uniform J8 icosahedral grid (655 362 nodes, 1 966 080
edges), 30 Boussinesq layers, synthetic stratified ocean state, FP64; one
step = one low-storage RK4 step of trend_ml (4 fused trend+update kernel
launches + 4 ghost fills).  DOF per step = layers * (2 * nodes + edges).
Inputs are resident in HBM before the timed region; W untimed warm-up steps,
then exactly K timed steps between two stream synchronisations.

Prints ONE JSON line.  `roofline` is the dominant kernel (the fused stage
kernel): algorithmic HBM bytes per launch / its mean launch time, both
measured here with HIP events on the kernel's stream.  `cpu_baseline` times
the CPU oracle (a one-core C port of the same algorithm) on a bounded sample
of the same workload (fewer layers, one step) on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "DOF updates/s J8 L30; free-surface solve ms @1e-9 rel. resid, learned vs MG"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_launch(layout, L: int) -> float:
    """Minimum HBM bytes of one fused stage launch, averaged over the four
    RK4 stages: per node-layer, read the stage state y (5 words), read the
    base state q (5 words; stage 1 has y = q), write the new state (5 words);
    per node, read the geometry once (len 3, pedlen 3, flux pedlen 3, area 6,
    hex_inv 1, coriolis 2, topography 1 = 19 words)."""
    nodes = 10 * layout.n * layout.n
    words_state = (10 + 15 + 15 + 15) / 4.0
    return 8.0 * (nodes * L * words_state + nodes * 19)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--layers", type=int, default=30)
    ap.add_argument("--cfl", type=float, default=0.5, help="dt = cfl * min edge length / sqrt(g * max depth)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (calibrates rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--cpu-layers", type=int, default=30, help="layers of the CPU-baseline sample")
    ap.add_argument("--cpu-steps", type=int, default=3, help="RK4 steps of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-free-surface", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-multilevel", action="store_true")
    ap.add_argument("--members", type=int, default=16, help="ensemble members of the batched-ensemble line")
    ap.add_argument("--fs-level", type=int, default=7, help="finest level of the free-surface hierarchy")
    ap.add_argument("--only-free-surface", action="store_true",
                    help="run only the free-surface line and print it (for profiling the solver)")
    ap.add_argument("--only-multilevel", action="store_true",
                    help="run only the multi-level line and print it (for profiling the hierarchy's kernels)")
    ap.add_argument("--ml-case", default="", help="with --only-multilevel: lo,hi,layers of one hierarchy (e.g. 5,8,30)")
    ap.add_argument("--ml-steps", type=int, default=3, help="steps timed per multi-level leg")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched right-hand-side line and print it (for profiling the batched solver)")
    ap.add_argument("--only-ensemble-drift", action="store_true",
                    help="run only config 5's drift line: --members J8 L30 members, --drift-steps split steps, FP64 "
                         "V-cycles against FP32-bulk stencils with learned cycles on the member-batched hierarchy")
    ap.add_argument("--drift-steps", type=int, default=100)
    ap.add_argument("--only-mlp", action="store_true",
                    help="run only the network's one-launch inference at 64 right-hand sides of J7 (163 840 rows)")
    ap.add_argument("--only-columns", action="store_true",
                    help="run only the column physics at --level/--layers: the vertical remap (PPR and remap1) and "
                         "the vertical diffusion (TKE closure and analytic), with their CPU baselines at J6")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding of the lozenge domains is a design note, not a code path", file=sys.stderr)
        sys.exit(2)
    if args.only_free_surface:
        print(json.dumps({"free_surface": free_surface(args.fs_level)}))
        return
    if args.only_multilevel:
        cs = ((tuple(int(v) for v in args.ml_case.split(",")),) if args.ml_case else ((5, 7, 20), (5, 8, 30)))
        print(json.dumps({"multilevel": multilevel(cs, nsteps=args.ml_steps)}))
        return
    if args.only_ensemble_drift:
        import wavetrisk_hydrostatic_amd as W
        from wavetrisk_hydrostatic_amd import cases
        grid = W.Grid(level=args.level, min_level=args.level, domain_level=2)
        topo, mass, temp, velo = cases.ocean(grid, args.layers, seed=2026)
        print(json.dumps({"ensemble_drift": ensemble_drift(grid, topo, mass, temp, velo, args.layers, args.members,
                                                           args.drift_steps)}))
        return
    if args.only_mlp:
        print(json.dumps({"mlp_forward_64_rhs": mlp_forward(64 * 10 * (2 ** args.fs_level // 8) ** 2)}))
        return
    if args.only_columns:
        print(json.dumps({"columns": columns(args.level, args.layers)}))
        return
    if args.only_batched:
        from wavetrisk_hydrostatic_amd import learned as LC
        grids = LC.hierarchy_grids(args.fs_level)
        npatch = 10 * (grids[-1].layout.n // 8) ** 2
        mlp = LC.load_network(npatch, os.path.join(os.path.dirname(LC.WEIGHTS), f"learned_J{args.fs_level}.npz"))
        print(json.dumps({"batched_64_rhs": free_surface_batched(grids, mlp, nrhs=64)}))
        mlp.close()
        return

    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import cases

    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"bench: copy calibration {args.copy_calibration} MiB read + written 3x", file=sys.stderr, flush=True)

    J, L = args.level, args.layers
    print(f"bench: building J{J} grid", file=sys.stderr, flush=True)
    grid = W.Grid(level=J, min_level=J, domain_level=2)
    lay = grid.layout
    topo, mass, temp, velo = cases.ocean(grid, L, seed=2026)
    own_e = lay.owned_edge_cells()
    args.dt = args.cfl * float(grid.array("len")[:, own_e].min()) / float(np.sqrt(9.80616 * np.max(-topo)))
    ctx = W.Context(grid, nlayers=L)
    ctx.set_topography(topo)
    ctx.upload_state(0, mass, temp, velo)
    print(f"bench: warmup {args.warmup} steps", file=sys.stderr, flush=True)
    if args.warmup:
        ctx.rk4(0, args.dt, args.warmup)
    ctx.synchronize()
    total_ms, kernel_ms = ctx.rk4_timed(0, args.dt, args.steps)
    ctx.synchronize()
    m_out, _, u_out = ctx.download_state(0)
    own = lay.owned_cells()
    if not (np.isfinite(m_out[:, own]).all() and np.isfinite(u_out[:, :, own_e]).all()):
        raise RuntimeError("non-finite state after the timed steps")

    dof_per_step = L * lay.n_dof_per_layer
    ms_per_step = total_ms / args.steps
    value = dof_per_step * args.steps / (total_ms / 1e3)
    alg_bytes = algorithmic_bytes_per_launch(lay, L)
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Oracle, geometry_of, num_threads

        nthr = num_threads()
        nst = args.cpu_steps
        print(f"bench: CPU baseline (oracle, {nthr} threads, J{J} L{args.cpu_layers}, {nst} RK4 steps)",
              file=sys.stderr, flush=True)
        Lc = args.cpu_layers
        t_c, m_c, tt_c, u_c = cases.ocean(grid, Lc, seed=2026)
        o = Oracle(lay, geometry_of(grid), t_c)
        m_c, tt_c, u_c = m_c.copy(), tt_c.copy(), u_c.copy()
        o.ghost_fill(m_c, tt_c, u_c)
        t0 = time.perf_counter()
        o.rk4(m_c, tt_c, u_c, args.dt, nst)
        secs = time.perf_counter() - t0
        cpu = {"value": nst * Lc * lay.n_dof_per_layer / secs, "unit": "DOF updates/s", "cores": nthr,
               "kind": "port",
               "sample": f"CPU oracle (C, -O2, OpenMP over the cell loops, {nthr} threads), {nst} RK4 steps of the "
                         f"same J{J} grid with {Lc} layers ({secs:.1f} s wall)"}

    bulk = fp32_bulk(grid, topo, mass, temp, velo, L, args.dt, args.warmup, args.steps, m_out, u_out,
                     dof_per_step, ms_per_step)
    shards = sharding_model(grid, ctx, L)
    ens = None if args.no_ensemble else ensemble(grid, topo, mass, temp, velo, L, args.dt, args.members)
    small = None if args.no_cpu else shallow_water_c1()
    ctx.synchronize()
    split = None if args.no_free_surface else split_step(grid, ctx, topo, mass, temp, velo, L)
    fs = None if args.no_free_surface else free_surface(args.fs_level, cpu=not args.no_cpu)
    ml = None if args.no_multilevel else multilevel()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded stratified ocean, analytic fields on the grid)",
        "config": {
            "workload": f"J{J} uniform icosahedral grid, L{L} Boussinesq layers, explicit RK4 of trend_ml "
                        f"(TRiSK), FP64, dt={args.dt:.1f}s (CFL {args.cfl}); no free-surface solve in this path",
            "level": J, "layers": L, "nodes": 10 * lay.n * lay.n + 2, "edges": 30 * lay.n * lay.n,
            "dof_per_step": dof_per_step,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(J, L),
            "kernel": "trisk_stage_kernel",
            "kernel_ms": kernel_ms,
            "algorithmic_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": cpu,
        "fp32_bulk": bulk,
        "sharding_model": shards,
        "shallow_water_J5": small,
        "ensemble": ens,
        "split_step": split,
        "free_surface": fs,
        "multilevel": ml,
    }
    print(json.dumps(out))


def measured_traffic(J: int, L: int, kernel: str = "trisk_stage"):
    """HBM bytes per launch of the stage kernel from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE + WRITE_SIZE per dispatch,
    scaled by a plain copy's readings), or None if it is absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_J{J}_L{L}_hbm.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def fp32_bulk(grid, topo, mass, temp, velo, L, dt, warmup, steps, m64, u64, dof_per_step, ms64):
    """The same J8 L30 RK4 steps with the FP32-bulk stencils
    (PRECISION_FP32_BULK): ms per step, DOF updates/s, mean stage-kernel
    launch, and the drift of the state from the FP64 run after the same
    warmup + timed steps (max-norm errors; surface height relative to the
    column depth)."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W

    print("bench: FP32-bulk steps", file=sys.stderr, flush=True)
    ctx = W.Context(grid, nlayers=L, nslots=1)
    ctx.set_topography(topo)
    ctx.upload_state(0, mass, temp, velo)
    ctx.set_precision(W.PRECISION_FP32_BULK)
    if warmup:
        ctx.rk4(0, dt, warmup)
    ctx.synchronize()
    total_ms, kernel_ms = ctx.rk4_timed(0, dt, steps)
    sh, tms = ctx.stage_shape()
    shape = {"chosen": {1: "layer pairs", 2: "edge tasks", 3: "one layer"}.get(sh, "auto"),
             "tuner_ms": {"layer pairs": tms[0], "edge tasks": tms[1], "one layer": tms[2]}}
    m, _, u = ctx.download_state(0)
    ctx.close()
    lay = grid.layout
    own, own_e = lay.owned_cells(), lay.owned_edge_cells()
    depth = m64[:, own].sum(0) / 1028.0
    eta_err = np.max(np.abs(m[:, own].sum(0) - m64[:, own].sum(0))) / 1028.0
    ms = total_ms / steps
    # algorithmic bytes of the mean stage: per node-layer and stage, 60 / 80 /
    # 80 / 100 B (FP64 state read, FP32 increment read and written, FP64 state
    # written in the last stage), and 80 B of geometry per node
    nodes = 10 * lay.n * lay.n
    alg = (0.25 * (60 + 80 + 80 + 100) * L + 80) * nodes
    achieved = alg / (kernel_ms / 1e3) / 1e9
    return {"ms_per_step": ms, "dof_updates_per_s": dof_per_step / (ms / 1e3), "speedup_vs_fp64": ms64 / ms,
            "stage_kernel_ms": kernel_ms,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(lay.level, L, "trisk_stage_mixed"),
                         "kernel": "trisk_stage_mixed_kernel", "algorithmic_bytes_per_launch": alg},
            "stage_shape": shape,
            "drift_vs_fp64": {"steps": warmup + steps,
                              "surface_height_rel_depth": eta_err / float(np.max(depth)),
                              "mass_rel": float(np.max(np.abs(m[:, own] - m64[:, own])) / np.max(np.abs(m64[:, own]))),
                              "velo_rel": float(np.max(np.abs(u[:, :, own_e] - u64[:, :, own_e])) /
                                                np.max(np.abs(u64[:, :, own_e])))},
            "note": "PV, fluxes, kinetic energy, divergence, Qperp in FP32; column scan, Bernoulli/Exner "
                    "gradients and state updates FP64; RK stages as FP32 increments over the FP64 state"}


def sharding_model(grid, ctx, L, xgmi_gbs=153.0):
    """Whole-lozenge shards of the J8 L30 state: ghost entries per layer that
    stay in a shard and that would cross to another GPU, the bytes one RK
    stage would send (all layers: mass, temp, velocity, FP64), the time of
    that at one xGMI link's bandwidth, and the measured time of the ghost
    fill's local and would-be-remote launches on this GPU."""
    print("bench: sharding model of the ghost fill", file=sys.stderr, flush=True)
    out = {"shard": "whole lozenges, lozenge lz on shard lz * nshards / 10", "xgmi_link_gbs": xgmi_gbs}
    for ns in (2, 5, 10):
        nl, nr, el, er = grid.shard_halo(ns)
        remote_bytes = L * (2 * nr + er) * 8
        ms_l, ms_r = ctx.ghost_fill_split_timed(0, ns, 20)
        out[str(ns)] = {"node_entries_local": nl, "node_entries_remote": nr, "edge_entries_local": el,
                        "edge_entries_remote": er, "remote_bytes_per_stage": remote_bytes,
                        "remote_ms_at_one_link": remote_bytes / (xgmi_gbs * 1e9) * 1e3,
                        "fill_ms_local": ms_l, "fill_ms_remote_entries": ms_r}
    return out


def shallow_water_c1(nsteps=100):
    """The reference's CPU-runnable case: single-layer shallow water (TC5,
    zonal flow over a mountain) on the uniform J5 grid, 100 RK4 steps at
    CFL 0.8; launch-bound on the GPU, so the steps are also timed replayed
    from a captured HIP graph.  The CPU oracle runs the same 100 steps."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import cases
    from oracle.oracle import Oracle, geometry_of, num_threads

    print("bench: J5 shallow water, eager and graph launches", file=sys.stderr, flush=True)
    g = W.Grid(level=5, min_level=5, domain_level=2)
    topo, m, t, u = cases.tc5(g, 1)
    own_e = g.layout.owned_edge_cells()
    depth = float(np.max(m.sum(0))) / 1028.0
    dt = 0.8 * float(g.array("len")[:, own_e].min()) / (float(np.sqrt(9.80616 * depth)) + 20.0)
    dof = g.layout.n_dof_per_layer
    out = {"workload": f"J5 TC5 single layer, {nsteps} RK4 steps, dt {dt:.1f} s (CFL 0.8)", "dof_per_step": dof}
    for name, prec, graph in (("fp64_eager", W.PRECISION_FP64, False), ("fp64_graph", W.PRECISION_FP64, True),
                              ("fp32_bulk_graph", W.PRECISION_FP32_BULK, True)):
        ctx = W.Context(g, nlayers=1, nslots=1)
        ctx.set_topography(topo)
        ctx.upload_state(0, m, t, u)
        ctx.set_precision(prec)
        ctx.set_graph(graph)
        ctx.rk4(0, dt, 4)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.rk4(0, dt, nsteps)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / nsteps
        ctx.close()
        out[name] = {"ms_per_step": ms, "dof_updates_per_s": dof / (ms / 1e3)}
    o = Oracle(g.layout, geometry_of(g), topo)
    mm, tt, uu = m.copy(), t.copy(), u.copy()
    o.ghost_fill(mm, tt, uu)
    t0 = time.perf_counter()
    o.rk4(mm, tt, uu, dt, nsteps)
    secs = time.perf_counter() - t0
    out["cpu_oracle"] = {"threads": num_threads(), "ms_per_step": secs * 1e3 / nsteps,
                         "dof_updates_per_s": dof * nsteps / secs}
    return out


def ensemble(grid, topo, mass, temp, velo, L, dt, members, nsteps=10, split_steps=2):
    """The batched ensemble (every layer kernel launched once for all
    members): members copies of the J8 L30 state, each with its own seeded
    bathymetry perturbation (+-50 m), RK4 steps in FP64 and in FP32-bulk; DOF
    updates/s of the whole ensemble, and the drift of each member's FP32-bulk
    state from its FP64 state after the same steps from the same start."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W

    print(f"bench: ensemble of {members} members", file=sys.stderr, flush=True)
    rng = np.random.default_rng(7)
    topos = np.stack([topo + 50.0 * np.tanh(rng.standard_normal()) * np.cos(np.arange(topo.size) * 1e-4 * (m + 1))
                      for m in range(members)])
    state = (np.tile(mass, (members, 1)), np.tile(temp, (members, 1)), np.tile(velo, (members, 1, 1)))
    res, split = {}, {}
    for prec in (W.PRECISION_FP64, W.PRECISION_FP32_BULK):
        ctx = W.Context(grid, nlayers=L, members=members, nslots=1)
        ctx.set_topography(topos)
        ctx.upload_state(0, *state)
        ctx.set_precision(prec)
        ctx.rk4(0, dt, 2)
        ctx.synchronize()
        total_ms, kernel_ms = ctx.rk4_timed(0, dt, nsteps)
        res[prec] = (total_ms, kernel_ms, ctx.download_state(0))
        # config 2's split-explicit step for every member: one trend launch for
        # all members, each member's free surface (BiCGSTAB to 1e-9) on its view
        ctx.upload_state(0, *state)
        ctx.fs_set_params(900.0, 0.6, 0.8)
        ctx.fs_upload(W.FS_ETA, state[0].reshape(members, L, -1).sum(1) / 1028.0 + topos)
        ctx.rk4_split(0, 1, 1e-9, 400)  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        its = [ctx.rk4_split(0, 1, 1e-9, 400) for _ in range(split_steps)]
        split[prec] = {"ms_per_step": (time.perf_counter() - t0) * 1e3 / split_steps,
                       "max_member_bicgstab_iterations": [int(i) for i in its]}
        ctx.close()
        split[prec]["hierarchy"] = ensemble_hierarchy_split(grid, topos, state, L, prec, split_steps)
    dof = members * L * grid.layout.n_dof_per_layer
    lay = grid.layout
    own, own_e = lay.owned_cells(), lay.owned_edge_cells()
    m64, _, u64 = res[W.PRECISION_FP64][2]
    m32, _, u32 = res[W.PRECISION_FP32_BULK][2]
    drift = []
    for k in range(members):
        sl = slice(k * L, (k + 1) * L)
        depth = m64[sl][:, own].sum(0) / 1028.0
        eta = np.max(np.abs(m32[sl][:, own].sum(0) - m64[sl][:, own].sum(0))) / 1028.0 / np.max(depth)
        du = np.max(np.abs(u32[sl][:, :, own_e] - u64[sl][:, :, own_e])) / np.max(np.abs(u64[sl][:, :, own_e]))
        drift.append({"surface_height_rel_depth": float(eta), "velo_rel": float(du)})
    t64, k64, _ = res[W.PRECISION_FP64]
    t32, k32, _ = res[W.PRECISION_FP32_BULK]
    return {"members": members, "ms_per_step": t64 / nsteps, "dof_updates_per_s": dof * nsteps / (t64 / 1e3),
            "stage_kernel_ms": k64,
            "fp32_bulk": {"ms_per_step": t32 / nsteps, "dof_updates_per_s": dof * nsteps / (t32 / 1e3),
                          "stage_kernel_ms": k32,
                          "drift_vs_fp64_per_member": {"steps": 2 + nsteps, "max_surface_height_rel_depth":
                                                       max(d["surface_height_rel_depth"] for d in drift),
                                                       "max_velo_rel": max(d["velo_rel"] for d in drift),
                                                       "members": drift}},
            "split_step": {"dt": 900.0, "rel_tol": 1e-9, "steps": split_steps,
                           "fp64": {**split[W.PRECISION_FP64], "dof_updates_per_s":
                                    dof / (split[W.PRECISION_FP64]["ms_per_step"] / 1e3)},
                           "fp32_bulk": {**split[W.PRECISION_FP32_BULK], "dof_updates_per_s":
                                         dof / (split[W.PRECISION_FP32_BULK]["ms_per_step"] / 1e3)},
                           "note": "host wall clock; the members' BiCGSTAB solves batched (one launch per step of "
                                   "the iteration for all members)"},
            "note": "one stage launch covers all members (grid y = member); same initial layers, per-member bathymetry"}


def ensemble_drift(grid, topo, mass, temp, velo, L, members, nsteps, dt=900.0):
    """Config 5 at its length: `members` copies of the J8 L30 ocean with their
    own bathymetry (as in the ensemble line), `nsteps` split steps with the
    members' free surfaces on the member-batched J(J-3)..J hierarchy to 1e-9:
    FP64 stencils with FP64 V-cycles, then FP32-bulk stencils with learned
    cycles (the shipped network; member-batched cycles smooth in FP64); wall
    ms per step of the whole ensemble, the largest member cycle count per
    step, and each member's drift of the FP32-bulk run from its FP64 run at
    the end (free surface relative to the column depth, mass, velocity)."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import learned as LC

    M = members
    print(f"bench: ensemble drift, {M} members, {nsteps} split steps", file=sys.stderr, flush=True)
    rng = np.random.default_rng(7)
    topos = np.stack([topo + 50.0 * np.tanh(rng.standard_normal()) * np.cos(np.arange(topo.size) * 1e-4 * (m + 1))
                      for m in range(M)])
    J = grid.layout.level
    grids = [W.Grid(level=l, min_level=J - 3, domain_level=2) for l in range(J - 3, J + 1)]
    tps = [topos]
    for k in range(len(grids) - 2, -1, -1):
        tps.insert(0, _subsample(grids[k + 1], grids[k], tps[0]))
    p = W.MgParams.defaults(coarse_tol=1e-2, fine_tol=1e-9, coarse_iter=400, max_vcycle=40)
    runs, final = {}, {}
    for name, prec, learned in (("fp64_vcycle", W.PRECISION_FP64, False),
                                ("fp32_bulk_learned", W.PRECISION_FP32_BULK, True)):
        ctxs = []
        for k, g in enumerate(grids):
            c = W.Context(g, nlayers=L if k == len(grids) - 1 else 1, members=M, nslots=1)
            c.set_topography(tps[k])
            c.fs_set_params(dt, 0.6, 0.8)
            ctxs.append(c)
        top = ctxs[-1]
        top.set_precision(prec)
        h = W.Hierarchy(ctxs)
        mlp = LC.load_network(h.npatches * M) if learned else None
        if learned and mlp is None:
            raise RuntimeError("no shipped network for the learned cycles")
        h.set_learned(mlp)
        top.upload_state(0, np.tile(mass, (M, 1)), np.tile(temp, (M, 1)), np.tile(velo, (M, 1, 1)))
        top.fs_upload(W.FS_ETA, np.tile(mass.sum(0) / 1028.0, (M, 1)) + topos)
        top.synchronize()
        meth = W.SOLVE_LEARNED if learned else W.SOLVE_VCYCLE
        t0 = time.perf_counter()
        cyc = [int(h.rk4_split(0, 1, meth, p)) for _ in range(nsteps)]
        top.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / nsteps
        runs[name] = {"ms_per_step": ms, "dof_updates_per_s": M * L * grid.layout.n_dof_per_layer / (ms / 1e3),
                      "max_member_cycles_min": min(cyc), "max_member_cycles_max": max(cyc),
                      "cycles_at_cap": sum(c >= p.max_vcycle for c in cyc)}
        m, _, u = top.download_state(0)
        final[name] = (m, u, top.fs_download(W.FS_ETA))
        h.set_learned(None)
        h.close()
        if mlp is not None:
            mlp.close()
        for c in ctxs:
            c.close()
    lay = grid.layout
    own, own_e = lay.owned_cells(), lay.owned_edge_cells()
    m64, u64, e64 = final["fp64_vcycle"]
    m32, u32, e32 = final["fp32_bulk_learned"]
    drift = []
    for k in range(M):
        sl = slice(k * L, (k + 1) * L)
        depth = np.max(m64[sl][:, own].sum(0) / 1028.0)
        drift.append({"surface_height_rel_depth": float(np.max(np.abs(e32[k][own] - e64[k][own])) / depth),
                      "mass_rel": float(np.max(np.abs(m32[sl][:, own] - m64[sl][:, own])) /
                                        np.max(np.abs(m64[sl][:, own]))),
                      "velo_rel": float(np.max(np.abs(u32[sl][:, :, own_e] - u64[sl][:, :, own_e])) /
                                        np.max(np.abs(u64[sl][:, :, own_e])))})
    return {"members": M, "steps": nsteps, "dt": dt, "levels": f"J{J - 3}..J{J}", "layers": L, **runs,
            "speedup_fp32_learned_vs_fp64": runs["fp64_vcycle"]["ms_per_step"] / runs["fp32_bulk_learned"]["ms_per_step"],
            "drift_vs_fp64": {"max_surface_height_rel_depth": max(d["surface_height_rel_depth"] for d in drift),
                              "max_mass_rel": max(d["mass_rel"] for d in drift),
                              "max_velo_rel": max(d["velo_rel"] for d in drift), "per_member": drift},
            "note": "host wall clock over all steps (no warm-up step: both runs start from the same state); "
                    "free surface to 1e-9 relative residual per member every step"}


def _subsample(fine, coarse, f):
    """a node field of the coarse level: the finer level's value at each
    coincident node of [0, n]^2 (halo cells are filled by the context)"""
    import numpy as np

    lz, i, j = coarse.layout.coords()
    m = (i >= 0) & (j >= 0) & (i <= coarse.layout.n) & (j <= coarse.layout.n)
    Sf = fine.layout.S
    out = np.zeros(f.shape[:-1] + (coarse.layout.ncells,))
    out[..., m] = f[..., lz[m] * Sf * Sf + (2 * j[m] + 2) * Sf + (2 * i[m] + 2)]
    return out


def ensemble_hierarchy_split(grid, topos, state, L, prec, nsteps, dt=900.0):
    """Config 5's split step with the members' free surfaces solved together
    on the J(J-3)..J hierarchy (member-batched cycles, coarse level to 1e-2
    per cycle, 1e-9 on the finest): plain V-cycles and learned cycles (the
    shipped network); ms per step of the whole ensemble."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import learned as LC

    M = topos.shape[0]
    J = grid.layout.level
    grids = [W.Grid(level=l, min_level=J - 3, domain_level=2) for l in range(J - 3, J + 1)]
    tps = [topos]
    for k in range(len(grids) - 2, -1, -1):
        tps.insert(0, _subsample(grids[k + 1], grids[k], tps[0]))
    ctxs = []
    for k, g in enumerate(grids):
        c = W.Context(g, nlayers=L if k == len(grids) - 1 else 1, members=M, nslots=1)
        c.set_topography(tps[k])
        c.fs_set_params(dt, 0.6, 0.8)
        ctxs.append(c)
    top = ctxs[-1]
    top.set_precision(prec)
    h = W.Hierarchy(ctxs)
    mlp = LC.load_network(h.npatches * M)
    p = W.MgParams.defaults(coarse_tol=1e-2, fine_tol=1e-9, coarse_iter=400, max_vcycle=40)
    out = {"levels": f"J{J - 3}..J{J}"}
    for name, meth in (("vcycle", W.SOLVE_VCYCLE), ("learned", W.SOLVE_LEARNED)):
        if meth == W.SOLVE_LEARNED:
            if mlp is None:
                continue
            h.set_learned(mlp)
        top.upload_state(0, *state)
        top.fs_upload(W.FS_ETA, state[0].reshape(M, L, -1).sum(1) / 1028.0 + topos)
        h.rk4_split(0, 1, meth, p)  # warm-up
        top.synchronize()
        t0 = time.perf_counter()
        cyc = [h.rk4_split(0, 1, meth, p) for _ in range(nsteps)]
        ms = (time.perf_counter() - t0) * 1e3 / nsteps
        out[name] = {"ms_per_step": ms, "dof_updates_per_s": M * L * grid.layout.n_dof_per_layer / (ms / 1e3),
                     "max_member_cycles": [int(c) for c in cyc]}
    h.set_learned(None)
    h.close()
    if mlp is not None:
        mlp.close()
    for c in ctxs:
        c.close()
    return out


def split_step(grid, ctx, topo, mass, temp, velo, L, dt=900.0, nsteps=5):
    """The ocean's split-explicit step (RK4_split: four trend stages at the
    baroclinic time step, then the implicit free surface solved by BiCGSTAB
    to 1e-9) on the same J8 L30 state: ms per step and DOF updates/s."""
    import numpy as np

    from wavetrisk_hydrostatic_amd import FS_ETA

    from wavetrisk_hydrostatic_amd import PRECISION_FP32_BULK, PRECISION_FP64

    print(f"bench: split-explicit steps, dt {dt} s", file=sys.stderr, flush=True)
    ctx.fs_set_params(dt, 0.6, 0.8)
    dof = L * grid.layout.n_dof_per_layer
    runs, final = {}, {}
    for name, prec in (("fp64", PRECISION_FP64), ("fp32_bulk", PRECISION_FP32_BULK)):
        ctx.set_precision(prec)
        ctx.upload_state(1, mass, temp, velo)
        ctx.fs_upload(FS_ETA, mass.sum(0) / 1028.0 + topo)
        ctx.rk4_split(1, 1, 1e-9, 400)  # warm-up (and, FP32-bulk, the stage autotuner)
        ctx.synchronize()
        t0 = time.perf_counter()
        its = [ctx.rk4_split(1, 1, 1e-9, 400) for _ in range(nsteps)]
        ms = (time.perf_counter() - t0) * 1e3 / nsteps
        runs[name] = {"ms_per_step": ms, "dof_updates_per_s": dof / (ms / 1e3),
                      "bicgstab_iterations": [int(i) for i in its]}
        final[name] = (ctx.download_state(1)[0], ctx.fs_download(FS_ETA))
    ctx.set_precision(PRECISION_FP64)
    # the same step with the free surface solved by V-cycles on the hierarchy
    # J(J-3)..J (the reference's multigrid setting): coarse levels one layer
    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import cases

    J = grid.layout.level
    mg_runs = {}
    grids = [W.Grid(level=l, min_level=J - 3, domain_level=2) for l in range(J - 3, J)]
    levels = []
    for g in grids:
        tp, _, _, _ = cases.ocean(g, 1, seed=3)
        c = W.Context(g, nlayers=1)
        c.set_topography(tp)
        c.fs_set_params(dt, 0.6, 0.8)
        levels.append(c)
    top = W.Context(W.Grid(level=J, min_level=J - 3, domain_level=2), nlayers=L)
    top.set_topography(topo)
    top.fs_set_params(dt, 0.6, 0.8)
    h = W.Hierarchy(levels + [top])
    p = W.MgParams.defaults(coarse_tol=1e-2, fine_tol=1e-9, coarse_iter=400, max_vcycle=40)
    from wavetrisk_hydrostatic_amd import learned as LC
    mlp = LC.load_network(h.npatches)
    runs_mg = [("fp64", PRECISION_FP64, W.SOLVE_VCYCLE), ("fp32_bulk", PRECISION_FP32_BULK, W.SOLVE_VCYCLE)]
    if mlp is not None:  # the north-star combination: FP32-bulk trends with learned cycles
        runs_mg += [("fp64_learned", PRECISION_FP64, W.SOLVE_LEARNED),
                    ("fp32_bulk_learned", PRECISION_FP32_BULK, W.SOLVE_LEARNED)]
        h.set_learned(mlp)
    for name, prec, meth in runs_mg:
        top.set_precision(prec)
        top.upload_state(0, mass, temp, velo)
        top.fs_upload(FS_ETA, mass.sum(0) / 1028.0 + topo)
        h.rk4_split(0, 1, meth, p)  # warm-up
        top.synchronize()
        t0 = time.perf_counter()
        cyc = [h.rk4_split(0, 1, meth, p) for _ in range(nsteps)]
        ms_mg = (time.perf_counter() - t0) * 1e3 / nsteps
        mg_runs[name] = {"ms_per_step": ms_mg, "dof_updates_per_s": dof / (ms_mg / 1e3),
                         "vcycles": [int(i) for i in cyc]}
    h.set_learned(None)
    if mlp is not None:
        mlp.close()
    h.close()
    for c in levels + [top]:
        c.close()
    own = grid.layout.owned_cells()
    depth = np.max(final["fp64"][0][:, own].sum(0) / 1028.0)
    drift = float(np.max(np.abs(final["fp32_bulk"][1][own] - final["fp64"][1][own])) / depth)
    return {"dt": dt, "theta1": 0.6, "theta2": 0.8, "rel_tol": 1e-9, **runs["fp64"],
            "fp32_bulk": {**runs["fp32_bulk"], "speedup_vs_fp64": runs["fp64"]["ms_per_step"] /
                          runs["fp32_bulk"]["ms_per_step"],
                          "surface_height_drift_rel_depth": drift, "steps_compared": nsteps + 1},
            "vcycle_hierarchy": {"levels": f"J{J - 3}..J{J}", "coarse_tol": 1e-2, **mg_runs},
            "note": "host wall clock over 5 steps after one warm-up step from the same state; BiCGSTAB queues 8 "
                    "gated iterations per host convergence test; FP32-bulk: FP32 trend stencils, FP64 stage "
                    "states, AXPYs and free surface"}


def multilevel(cases_=((5, 7, 20), (5, 8, 30)), dt=300.0, nsteps=3):
    """The multi-level step on a full hierarchy (every node active, as the
    reference runs a uniform hierarchy with zero thresholds): explicit RK4
    with trend_ml and WT_after_step over levels, and RK4_split with the FMG
    free surface to 1e-9 on every level -- config 2's J5..J7 L20 ocean and
    config 5's member shape J5..J8 L30.  Host wall clock over nsteps steps
    after one warm-up step; DOF counted on the finest level and on all."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import cases

    out = {}
    for lo, hi, L in cases_:
        print(f"bench: multi-level steps J{lo}..J{hi} L{L}", file=sys.stderr, flush=True)
        grids = [W.Grid(level=l, min_level=lo, domain_level=2) for l in range(lo, hi + 1)]
        ctxs, st, topos = [], [], []
        for g in grids:
            topo, m, t, u = cases.ocean(g, L, seed=21)
            c = W.Context(g, nlayers=L, nslots=2)
            c.set_topography(topo)
            c.fs_set_params(dt, 0.6, 0.8)
            ctxs.append(c)
            st.append((m, t, u))
            topos.append(topo)
        h = W.Hierarchy(ctxs)
        ml = W.MultiLevel(ctxs)
        ml.upload(0, [s[0] for s in st], [s[1] for s in st], [s[2] for s in st])
        ml.forward(0)
        start = ml.download(0)
        dof_fine = L * grids[-1].layout.n_dof_per_layer
        dof_all = sum(L * g.layout.n_dof_per_layer for g in grids)
        # explicit RK4 over the hierarchy
        ml.upload(0, *start)
        ml.rk4(0, dt, 1)
        ctxs[-1].synchronize()
        t0 = time.perf_counter()
        ml.rk4(0, dt, nsteps)
        ms_rk4 = (time.perf_counter() - t0) * 1e3 / nsteps
        # RK4_split with FMG to 1e-9
        p = W.MgParams.defaults(coarse_tol=1e-9, fine_tol=1e-9, coarse_iter=400, max_vcycle=50)
        ml.upload(0, *start)
        for i, c in enumerate(ctxs):
            c.fs_upload(W.FS_ETA, start[0][i].sum(0) / 1028.0 + topos[i])
        ml.rk4_split(h, 0, 1, p)
        ctxs[-1].synchronize()
        t0 = time.perf_counter()
        its = [ml.rk4_split(h, 0, 1, p) for _ in range(nsteps)]
        ms_split = (time.perf_counter() - t0) * 1e3 / nsteps
        m, _, u = ml.download(0)
        finite = bool(all(np.isfinite(a).all() for a in m) and all(np.isfinite(a).all() for a in u))
        ml.close()
        h.close()
        for c in ctxs:
            c.close()
        out[f"J{lo}..J{hi}_L{L}"] = {
            "dof_per_step_finest": dof_fine, "dof_per_step_all_levels": dof_all,
            "rk4": {"ms_per_step": ms_rk4, "dof_updates_per_s_finest": dof_fine / (ms_rk4 / 1e3)},
            "rk4_split_fmg": {"ms_per_step": ms_split, "dof_updates_per_s_finest": dof_fine / (ms_split / 1e3),
                              "fmg_vcycles_summed_over_levels": [int(i) for i in its], "finite": finite}}
    out["note"] = (f"dt {dt} s, theta1 0.6, theta2 0.8, FMG to 1e-9 on every level (the reference's parameters "
                   f"otherwise), FP64; host wall clock over {nsteps} steps after one warm-up step")
    return out


def mlp_forward(rows: int, params=None):
    """The network's one-launch inference (mlp_infer_kernel) on `rows` rows of
    seeded features: mean ms of 20 launches (HIP events), BF16 FLOP rate and
    its fraction of MI355X's dense BF16 peak"""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W

    big = W.MLP(max_batch=rows, seed=1)
    if params is not None:
        big.params = params
    big.forward(np.random.default_rng(0).standard_normal((rows, W.MLP_IN)).astype(np.float32))
    ms = big.forward_timed(rows, 20)
    flops = 2.0 * (W.MLP_IN * W.MLP_HIDDEN + W.MLP_HIDDEN * W.MLP_HIDDEN + W.MLP_HIDDEN * W.MLP_OUT) * rows
    big.close()
    # bytes: BF16 inputs read, FP32 outputs written
    hbm = rows * (2 * W.MLP_IN + 4 * W.MLP_OUT)
    return {"rows": rows, "ms_per_forward": ms, "tflops": flops / (ms * 1e-3) / 1e12,
            "frac_of_bf16_dense_peak": flops / (ms * 1e-3) / 2.5e15,
            "hbm_gbs": hbm / (ms * 1e-3) / 1e9, "hbm_bound_ms": hbm / 8e12 * 1e3}


def columns(level: int, L: int, calls: int = 10):
    """The column physics after a step: wtg_remap (remapPPR and remap1 onto
    Chebyshev hybrid coordinates) and wtg_vertical_diffusion (TKE closure and
    analytic coefficients) on the ocean at `level` with L layers: mean wall
    ms per call (each call ghost-fills before and after and synchronizes) over
    `calls` calls after one warm-up, the algorithmic HBM rate (remap: m, t, 3 u
    read and written plus the old-mass copy, 112 B per node-layer; vertical
    diffusion: two passes reading m, t, 3 u, Kt, Kv, TKE and the forcing and
    writing t, 3 u, Kt, Kv, TKE, 176 B per node-layer) and the oracle's time
    on the same physics at J6 (one thread), scaled per node-layer"""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import cases
    from oracle.oracle import Oracle, geometry_of

    def setup(lev):
        g = W.Grid(level=lev, min_level=lev, domain_level=2)
        topo, m, t, u = cases.ocean(g, L, seed=11)
        rng = np.random.default_rng(3)
        NC = g.layout.ncells
        forcing = dict(bot=rng.uniform(-3e-7, 3e-7, NC), top=rng.uniform(-1e-6, 1e-6, NC),
                       tau=rng.uniform(0.02, 0.2, NC), wind=rng.uniform(-2e-4, 2e-4, (3, NC)))
        coef = (np.full((L + 1, NC), 1.2e-5), np.full((L + 1, NC), 1.2e-4), rng.uniform(1e-6, 3e-5, (L, NC)))
        return g, topo, m, t, u, forcing, coef

    k = np.arange(L + 1)
    b = (1.0 + np.cos(k / L * np.pi)) / 2.0
    a = 1.0 - b
    g, topo, m, t, u, forcing, coef = setup(level)
    nl = 10 * g.layout.n ** 2 * L  # node-layers
    ctx = W.Context(g, nlayers=L)
    ctx.set_topography(topo)
    out = {"workload": f"J{level} L{L} ocean", "calls": calls}

    def timed(fn):
        ctx.upload_state(0, m, t, u)
        fn()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(calls):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t0) / calls * 1e3

    for name, scheme in (("remap_ppr", W.REMAP_PPR), ("remap_linear", W.REMAP_PIECEWISE_LINEAR)):
        ms = timed(lambda: ctx.remap(0, scheme, a, b))
        out[name] = {"ms_per_call": ms, "alg_gbs": nl * 112 / (ms * 1e-3) / 1e9}
    for name, tk in (("vdiff_tke", 1), ("vdiff_analytic", 0)):
        ctx.set_vdiff(tke_closure=tk, bottom_friction=4e-4)
        ctx.vdiff_set_coefficients(*coef)
        ctx.vdiff_set_forcing(forcing["bot"], forcing["top"], forcing["tau"], forcing["wind"])
        ms = timed(lambda: ctx.vertical_diffusion(0, 900.0, 2))
        out[name] = {"ms_per_call": ms, "alg_gbs": nl * 176 / (ms * 1e-3) / 1e9}
    ctx.close()
    # CPU baseline: the oracle at J6 (one thread), per node-layer scaled to the GPU workload
    g6, topo6, m6, t6, u6, f6, c6 = setup(6)
    nl6 = 10 * g6.layout.n ** 2 * L
    o = Oracle(g6.layout, geometry_of(g6), topo6)
    st = (m6.copy(), t6.copy(), u6.copy())
    o.ghost_fill(*st)
    t0 = time.perf_counter()
    o.remap(*st, W.REMAP_PPR, a, b)
    ppr6 = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.vertical_diffusion(*st, 900.0, 2, f6, *c6, tke_closure=1, bottom_friction=4e-4)
    vd6 = time.perf_counter() - t0
    out["cpu_baseline"] = {"kind": "port", "cores": 1, "sample": f"the C oracle at J6 L{L} (one thread), scaled "
                                                                 f"by node-layers to J{level}",
                           "remap_ppr_ms": ppr6 * nl / nl6 * 1e3, "vdiff_tke_ms": vd6 * nl / nl6 * 1e3}
    return out


def free_surface_cpu(prob, dt: float = 1800.0):
    """CPU baseline of the solve: the oracle's FMG (the reference's method:
    FMG with BiCGSTAB on the coarsest level, tolerances 1e-9) on the same
    draw and hierarchy, one thread"""
    import numpy as np

    from oracle.oracle import MgOracle

    mg = MgOracle(prob.grids, prob.topos, prob.eta_olds, dt=dt, coarse_tol=1e-9, fine_tol=1e-9, coarse_iter=400,
                  max_vcycle=40)
    zero = [np.zeros(g.layout.ncells) for g in prob.grids]
    t0 = time.perf_counter()
    _, it, res = mg.fmg(zero, prob.rhs)
    secs = time.perf_counter() - t0
    return {"value": secs * 1e3, "unit": "ms per FMG solve to 1e-9", "cores": 1, "kind": "port",
            "sample": "the C oracle's FMG (wto_fmg, one thread) on the same J5..J7 draw",
            "iterations": [int(v) for v in it], "rel_resid": float(res[-1])}


def free_surface(fine: int, cpu: bool = True):
    """The free-surface solve to 1e-9 relative residual on a uniform
    hierarchy J(fine-2)..J(fine) of a seeded ocean draw (dt = 1800 s): plain
    V-cycles from zero, the same cycles with the learned correction (the
    shipped network if present, else one trained here on harvested pairs),
    and the reference's FMG; cycles, final FP64 residual and wall-clock ms
    (median of 3, host-driven including its residual read-backs)."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import SOLVE_FMG, SOLVE_LEARNED, SOLVE_VCYCLE, MgParams
    from wavetrisk_hydrostatic_amd import learned as LC

    print(f"bench: free surface, hierarchy J{fine - 2}..J{fine}", file=sys.stderr, flush=True)
    grids = LC.hierarchy_grids(fine)
    prob = LC.Problem(grids, seed=4242)
    npatch = prob.h.npatches
    mlp = LC.load_network(npatch, os.path.join(os.path.dirname(LC.WEIGHTS), f"learned_J{fine}.npz"))
    source = "shipped weights"
    if mlp is None:
        print("bench: no shipped network, training one (6 draws, 1500 steps)", file=sys.stderr, flush=True)
        X, T = [], []
        for s in range(6):
            pr = LC.Problem(grids, seed=100 + s)
            x, t = LC.harvest(pr, ncycles=5)
            pr.close()
            X.append(x)
            T.append(t)
        mlp, _ = LC.train_network(np.concatenate(X), np.concatenate(T), steps=1500, batch=1024, max_batch=npatch)
        source = "trained in this run on 6 draws"

    def timed(method, p, learned=False):
        prob.h.set_learned(mlp if learned else None)
        runs = [prob.solve(method, p) for _ in range(4)][1:]
        prob.h.set_learned(None)
        ms = sorted(r[2] for r in runs)[1]
        it, res, _ = runs[-1]
        return {"iterations": [int(v) for v in it], "rel_resid": float(res[-1]), "ms": ms}

    # the network at config 4's batch: one row per patch for 64 right-hand sides
    mlp_line = mlp_forward(64 * npatch, mlp.params)

    pv = LC.vcycle_params(1e-9, 40)
    pv1 = LC.vcycle_params(1e-9, 40)
    pv1.learned_cycles = 1  # the network's correction before the first cycle only
    # two more seeded draws: plain V-cycles, learned cycles, learned first cycle
    draws = {}
    for seed in (7, 99):
        pr = LC.Problem(grids, seed=seed)
        row = {}
        for name, meth, par, lrn in (("vcycle", SOLVE_VCYCLE, pv, False), ("learned", SOLVE_LEARNED, pv, True),
                                     ("learned_first_cycle", SOLVE_LEARNED, pv1, True)):
            pr.h.set_learned(mlp if lrn else None)
            runs = [pr.solve(meth, par) for _ in range(4)][1:]
            row[name] = {"cycles": int(runs[-1][0][-1]), "rel_resid": float(runs[-1][1][-1]),
                         "ms": sorted(r[2] for r in runs)[1]}
        pr.h.set_learned(None)
        pr.close()
        draws[f"seed_{seed}"] = row
    # V-cycles and learned cycles with FP32 smoothing (FP64 residuals and update)
    prob.h.set_precision(W.PRECISION_FP32_BULK)
    fp32 = {"vcycle": timed(SOLVE_VCYCLE, pv), "learned": timed(SOLVE_LEARNED, pv, learned=True),
            "note": "each cycle a defect correction: FP64 residual, correction smoothed and transferred in FP32 "
                    "fields, coarsest BiCGSTAB FP64"}
    prob.h.set_precision(W.PRECISION_FP64)
    out = {"workload": f"elliptic_lo on J{fine - 2}..J{fine} (finest {10 * 4 ** fine + 2} nodes), dt 1800 s, "
                       "seeded ocean draw 4242, tolerance 1e-9",
           "vcycle": timed(SOLVE_VCYCLE, pv),
           "learned": timed(SOLVE_LEARNED, pv, learned=True),
           "fmg": timed(SOLVE_FMG, MgParams.defaults(coarse_tol=1e-9, fine_tol=1e-9, coarse_iter=400, max_vcycle=40)),
           "network": source,
           "learned_first_cycle": timed(SOLVE_LEARNED, pv1, learned=True),
           "draws": draws,
           "fp32_smoothing": fp32,
           "mlp_forward_64_rhs": mlp_line}
    if cpu:
        out["fmg_cpu_baseline"] = free_surface_cpu(prob)
        out["fmg_cpu_baseline"]["gpu_speedup"] = out["fmg_cpu_baseline"]["value"] / out["fmg"]["ms"]
    out["batched_64_rhs"] = free_surface_batched(grids, mlp, nrhs=64)
    prob.close()
    mlp.close()
    return out


def free_surface_batched(grids, mlp, nrhs=64):
    """Config 4's batch: nrhs right-hand sides (seeded draws, each with its
    own bathymetry and old free surface) on the same hierarchy, solved to
    1e-9 in the same launches (an ensemble hierarchy: one launch per operator
    of the V-cycle for every right-hand side, one network inference for all
    their patches); plain V-cycles against learned cycles (the same network,
    its step before every cycle and before the first only).  Wall-clock ms
    for the whole batch (median of 3 after a warm-up) and per right-hand side;
    cycles per right-hand side."""
    import numpy as np

    import wavetrisk_hydrostatic_amd as W
    from wavetrisk_hydrostatic_amd import learned as LC

    print(f"bench: free surface, {nrhs} right-hand sides batched", file=sys.stderr, flush=True)
    prob = LC.Problem(grids, [5000 + k for k in range(nrhs)])
    big = W.MLP(max_batch=prob.h.npatches * nrhs, seed=1)
    big.params = mlp.params
    pv = LC.vcycle_params(1e-9, 40)
    pv1 = LC.vcycle_params(1e-9, 40)
    pv1.learned_cycles = 1
    out = {"rhs": nrhs}
    for name, meth, par, lrn in (("vcycle", W.SOLVE_VCYCLE, pv, False), ("learned", W.SOLVE_LEARNED, pv, True),
                                 ("learned_first_cycle", W.SOLVE_LEARNED, pv1, True)):
        prob.h.set_learned(big if lrn else None)
        runs = [prob.solve(meth, par) for _ in range(4)][1:]
        cyc, res = prob.h.member_results()
        ms = sorted(r[2] for r in runs)[1]
        out[name] = {"ms_batch": ms, "ms_per_rhs": ms / nrhs, "cycles_mean": float(np.mean(cyc)),
                     "cycles_max": int(np.max(cyc)), "cycles": [int(c) for c in cyc],
                     "max_rel_resid": float(np.max(res))}
    prob.h.set_learned(None)
    prob.close()
    big.close()
    out["note"] = "right-hand sides as ensemble members of one hierarchy; a converged member is left alone"
    return out


if __name__ == "__main__":
    main()
