#!/usr/bin/env python3
"""Benchmark of the PIC/FLIP step on one MI355X.

Workload (default): 256^3 MAC grid, solid walls, water layer y < 0.4 seeded
with 8 jittered particles per cell, rigid sphere r = 0.08 (density 500) released
at rest at (0.5, 0.8, 0.5), two-way coupled; FLIP 0.95 (PIC weight 0.05),
dt = min(CFL 1, 1/120); pressure by MGPCG to 1e-6 relative residual with FP64
dots and FP64 refinement.  A "step" is one full frameStep (label + sort, P2G,
forces, RHS + pressure solve, projection, body, G2P + advection) with every
input resident in HBM.

Prints one JSON line (see the keys below).  The CPU baseline is the oracle
(oracle/_build/libfs_oracle.so, OpenMP) timed on a bounded sample of the same
workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s at 256³ FLIP; pressure-solve ms to 1e-6 residual, learned vs plain"
HBM_PEAK_GBS = 8000.0


def build_scene(name):
    from fluidprjct_amd import scenes
    if name == "c3":
        N = 256
        lab = scenes.water_column(N, 0.4)
        r = 0.08
        sdf, origin, cs = scenes.sphere_sdf(r, 1.0 / N)
        body = dict(sdf=sdf, origin=origin, cell_size=cs, pos=(0.5, 0.8, 0.5), mass=scenes.sphere_mass(r, 500.0))
        desc = ("256^3 water column y<0.4, 8 ppc, rigid sphere r=0.08 rho=500 at (0.5,0.8,0.5), "
                "FLIP 0.95, dt=min(CFL 1,1/120), MGPCG 1e-6")
        return N, lab, body, 8, 3, desc
    if name == "c1":
        N = 128
        lab = scenes.dam_break_2d(N, (0.25, 0.5))
        return N, lab, None, 4, 1, "128^2 2D dam break x<=0.25 y<=0.5 (one-cell-deep slab), 4 ppc, FLIP 0.95, CFL 1"
    if name == "c2":
        N = 128
        lab = scenes.dam_break(N, (0.25, 0.5, 1.0))
        return N, lab, None, 8, 2, "128^3 dam break x<=0.25 y<=0.5, 8 ppc, FLIP 0.95, CFL 1, MGPCG 1e-6"
    if name == "c5":
        N = 512
        lab = scenes.shallow_basin(N, seed=5)
        return N, lab, None, 8, 5, "512^3 shallow basin over 32 Gaussians, y<0.3 tilted 0.05, 8 ppc"
    raise SystemExit(f"unknown config {name}")


def make_solver(lib, N, lab, body, ppc, seed, tol=1e-6, flags=0, step_flags=0):
    s = lib.solver(dt=1.0 / 120, dx=1.0 / N, pic_weight=0.05, cfl=1.0, solver_tol=tol, solver_max_iter=400,
                   particles_per_cell=ppc, seed=seed, solver_flags=flags, step_flags=step_flags)
    s.set_labels(lab)
    if body is not None:
        s.set_body_sdf(body["sdf"], body["origin"], body["cell_size"], body["pos"], body["mass"])
    s.seed_particles(ppc, seed)
    return s


KERNEL_NAMES = {"g2p_advect": "k_g2p_advect_tile", "p2g": "k_p2g_tile", "project": "k_project"}


def measured_traffic(kernel, config):
    """HBM bytes per launch of kernel from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{config}_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def stage_delta(a, b):
    return {k: b[k] - a[k] for k in a}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def face_gradient_norm(dp, lab):
    """L2 norm of the pressure differences across the faces that the projection
    updates with p (both cells FLUID or AIR, FluidSolver3D.cu:2072-2336)."""
    ok = (lab == 1) | (lab == 2)
    tot = 0.0
    for ax in (2, 1, 0):
        a = [slice(None)] * 3
        b = [slice(None)] * 3
        a[ax], b[ax] = slice(1, None), slice(None, -1)
        m = ok[tuple(a)] & ok[tuple(b)]
        d = (dp[tuple(a)] - dp[tuple(b)])[m]
        tot += float(d @ d)
    return tot ** 0.5


def projected_velocity_norm(s, lab):
    from fluidprjct_amd._abi import FS_GRID_U, FS_GRID_V, FS_GRID_W
    tot = 0.0
    fl = lab == 1
    for g, ax in ((FS_GRID_U, 2), (FS_GRID_V, 1), (FS_GRID_W, 0)):
        f = s.get_grid(g).astype(np.float64)
        a = [slice(None)] * 3
        b = [slice(None)] * 3
        a[ax], b[ax] = slice(1, None), slice(None, -1)
        # faces with a FLUID cell on either side
        m = np.zeros(f.shape, bool)
        m[tuple(a)] |= fl
        m[tuple(b)] |= fl
        tot += float((f[m] ** 2).sum())
    return tot ** 0.5


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200,
                    help="timed steps from the workload's initial state (the 256^3 sphere drop is 200 steps)")
    ap.add_argument("--warmup", type=int, default=3, help="untimed steps on a separate handle first")
    ap.add_argument("--config", default="c3")
    ap.add_argument("--no-phases", action="store_true",
                    help="skip the per-step run that splits the workload into its phases (before/after impact)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch the inner PCG iterations eagerly")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--deterministic", action="store_true",
                    help="FS_STEP_DETERMINISTIC: particles of a cell in id order (bit-reproducible runs)")
    ap.add_argument("--voxel-coupling", action="store_true",
                    help="couple the body as the reference does (BODY cells at cell centres only) instead of "
                         "through its SDF face fractions (FS_STEP_FACE_FRACTIONS), which the 256^3 "
                         "sphere-drop workload names")
    ap.add_argument("--no-other-coupling", action="store_true",
                    help="skip the leg that times the same workload with the other body coupling")
    ap.add_argument("--learned-steps", type=int, default=5)
    ap.add_argument("--learned-params", default="", help="parameters of the learned preconditioner (.npy)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a float32 y = x + 1 kernel over MIB MiB (read MIB, write MIB): the "
                         "reference for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the PIC/FLIP step is not sharded across GPUs in this build "
              "(z-slab decomposition is a design note, not code).", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.zeros(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n}))
        return
    from fluidprjct_amd import load_library
    from fluidprjct_amd._abi import (FS_GRID_LABELS, FS_GRID_P64, FS_GRID_RHS, FS_STEP_DETERMINISTIC,
                                     FS_STEP_FACE_FRACTIONS)
    hip = load_library()

    N, lab, body, ppc, seed, desc = build_scene(args.config)
    log(f"{args.config}: scene {lab.shape}")
    step_flags = ((FS_STEP_DETERMINISTIC if args.deterministic else 0) |
                  (FS_STEP_FACE_FRACTIONS if not args.voxel_coupling and body is not None else 0))
    solver_flags = 1 if args.no_graph else 0

    # warmup: the same workload's first steps on a handle of its own (code
    # objects loaded, pressure graphs of the workload's box shapes captured
    # once per process), so that the timed run starts from the released state
    if args.warmup > 0:
        w = make_solver(hip, N, lab, body, ppc, seed, flags=solver_flags, step_flags=step_flags)
        log(f"{w.num_particles} particles; {args.warmup} warmup steps on a separate handle")
        w.step(args.warmup)
        torch.cuda.synchronize()
        del w
        torch.cuda.empty_cache()

    # the timed run: exactly K steps of the workload from its initial state
    s = make_solver(hip, N, lab, body, ppc, seed, flags=solver_flags, step_flags=step_flags)
    Np = s.num_particles
    torch.cuda.synchronize()
    t_stage0 = s.stage_times()
    log(f"{args.steps} timed steps from the initial state")
    t0 = time.perf_counter()
    rc_steps = s.lib.fs_step(s.h, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    not_converged = rc_steps == -4
    if rc_steps not in (0, -4):
        s.fl.check(rc_steps)
    t_stage = stage_delta(t_stage0, s.stage_times())
    K = args.steps
    st = s.solve_stats()
    simt, dt_last, nsteps = s.clock()

    D, H, W = lab.shape
    C = W * H * D
    F = ((W + 1) * H * D + W * (H + 1) * D + W * H * (D + 1)) // 3   # faces per component (mean)
    labs = s.get_grid(FS_GRID_LABELS)
    nf = int((labs == 1).sum())
    ms = {k: v / K for k, v in t_stage.items()}
    # algorithmic bytes per launch (SoA FP32 particles 24 B, uint8 labels, FP32 faces)
    alg = {
        "g2p_advect": 48 * Np + 36 * F + C,       # read pos+vel, write pos+vel, read u,v,w,uSaved.. (6x4 B x F), labels
        "p2g": 24 * Np + 24 * F,                  # read pos+vel, write u,v,w and uSaved,vSaved,wSaved
        "project": 24 * F + 5 * C,
    }
    dom = max(alg, key=lambda k: ms[k])
    achieved = alg[dom] / (ms[dom] * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": K / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": args.config + ": " + desc + (", deterministic cell order" if args.deterministic else "")
                   + (", SDF face fractions" if step_flags & FS_STEP_FACE_FRACTIONS else ""),
                   "grid": [W, H, D], "particles": Np,
                   "fluid_cells_last": nf, "precision": "FP32 transfers/stencil, FP64 dots + refinement",
                   "timed": f"steps 1..{K} of the workload from its initial state (one fs_step call); the "
                            f"{args.warmup} warmup steps ran on a separate handle"},
        "pressure_solve_ms_plain": ms["pressure"],
        "pressure_iters_last": st.iterations,
        "pressure_outer_last": st.outer,
        "pressure_rel_residual_last": st.rel_residual,
        "pressure_not_converged": bool(not_converged),
        "stage_ms": {k: round(v, 4) for k, v in ms.items()},
        "particles_per_s_transfers": Np / ((ms["p2g"] + ms["g2p_advect"]) * 1e-3),
        "dt_last": dt_last,
        "sim_time": simt,
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(KERNEL_NAMES[dom], args.config), "alg_bytes": alg[dom],
                     "avg_launch_ms": ms[dom],
                     "note": "alg_bytes per launch / the stage's mean HIP-event time over the timed steps"},
    }
    if body is not None:
        result["body_y_final"] = s.body_state().pos[1]

    # one more step whose system (labels, right-hand side, body pose) the FP64
    # check below solves again
    sys_body = s.body_state() if body is not None else None
    s.step(1)
    sys_lab = s.get_grid(FS_GRID_LABELS)
    sys_b = s.get_grid(FS_GRID_RHS)
    sys_p = s.get_grid(FS_GRID_P64)
    sys_dt = s.clock()[1]
    u_norm = projected_velocity_norm(s, sys_lab)

    # learned vs plain, after the workload: the same run continued for a few
    # steps with each preconditioner, then one system solved by both from zero
    if not args.no_learned:
        from fluidprjct_amd._abi import FS_PRECOND_HYBRID, FS_PRECOND_LEARNED, FS_PRECOND_MG
        from fluidprjct_amd.learned import DEFAULT_PARAMS, HYBRID_PARAMS, load_params
        params = args.learned_params or DEFAULT_PARAMS
        modes = [("plain", FS_PRECOND_MG, None), ("learned", FS_PRECOND_LEARNED, params)]
        if os.path.exists(HYBRID_PARAMS):
            modes.append(("hybrid", FS_PRECOND_HYBRID, HYBRID_PARAMS))
        legs = {}
        for name, pc, prm in modes:
            if prm is not None:
                s.set_learned_params(load_params(prm))
            s.set_preconditioner(pc, 5)
            if name != "plain":
                s.step(1)
            torch.cuda.synchronize()
            t0s = s.stage_times()
            fb, iters = 0, []
            for _ in range(args.learned_steps):
                s.step(1)
                stl = s.solve_stats()
                fb += stl.fallback
                iters.append(stl.iterations)
            torch.cuda.synchronize()
            tl = stage_delta(t0s, s.stage_times())
            legs[name] = {"pressure_ms": tl["pressure"] / args.learned_steps, "iterations": iters, "fallbacks": fb}
            log(f"{name} steps: {legs[name]}")
        same = {}
        for name, pc, prm in modes:
            if prm is not None:
                s.set_learned_params(load_params(prm))
            s.set_preconditioner(pc, 5)
            s.solve_pressure_system(sys_lab, sys_b, 1e-6)
            _, sts = s.solve_pressure_system(sys_lab, sys_b, 1e-6)
            same[name] = {"ms": sts.solve_ms, "iterations": sts.iterations,
                          "fallback": sts.fallback, "rel_residual": sts.rel_residual}
        result["pressure_solve_ms_learned"] = legs["learned"]["pressure_ms"]
        result["pressure_solve_ms_hybrid"] = legs["hybrid"]["pressure_ms"] if "hybrid" in legs else None
        result["learned"] = {
            "params": {name: os.path.relpath(prm, ROOT) for name, _, prm in modes if prm},
            "steps": legs, "same_system": same,
            "note": f"steps: the workload continued after step {K + 1} for {args.learned_steps} steps with "
                    f"each preconditioner (pressure stage mean; learned: the network alone in flexible PCG, hybrid: "
                    f"a V-cycle corrected by the network, both with the multigrid fallback after 5 stalled "
                    f"iterations); same_system: the system of step {K + 1} "
                    f"solved from p = 0 by both (the reference's matrix; the second solve's device time, without the host copies)"}
    del s
    torch.cuda.empty_cache()

    # phases of the same workload on a fresh handle, one fs_step per step:
    # the resting column before the sphere reaches the water and after it
    if not args.no_phases:
        ph = make_solver(hip, N, lab, body, ppc, seed, flags=solver_flags, step_flags=step_flags)
        step_ms, submerged = [], []
        log(f"phases: {K} steps one at a time")
        for _ in range(K):
            t = time.perf_counter()
            ph.step(1)
            step_ms.append(1e3 * (time.perf_counter() - t))
            if body is not None:
                submerged.append(ph.body_state().submerged_volume)
        phases = {"steps_per_s": K / (sum(step_ms) * 1e-3),
                  "steps_4_23_per_s": 20 / (sum(step_ms[3:23]) * 1e-3) if K >= 23 else None,
                  "note": "host wall clock per step (each fs_step returns synchronised; the body state is read "
                          "after every step outside the clock)"}
        if submerged:
            hit = next((i for i, v in enumerate(submerged) if v > 0), None)
            phases["impact_step"] = hit
            if hit:
                pre, post = step_ms[1:hit], step_ms[hit:]
                phases["pre_impact_steps_per_s"] = len(pre) / (sum(pre) * 1e-3) if pre else None
                phases["post_impact_steps_per_s"] = len(post) / (sum(post) * 1e-3) if post else None
        result["phases"] = phases
        del ph
        torch.cuda.empty_cache()

    # the same workload with the other body coupling (the reference's voxel
    # coupling when the line is quoted with face fractions, and vice versa),
    # K steps from the initial state on a fresh handle
    if body is not None and not args.no_other_coupling:
        other = step_flags ^ FS_STEP_FACE_FRACTIONS
        log("other body coupling: " + ("SDF face fractions" if other & FS_STEP_FACE_FRACTIONS else "voxels"))
        s2 = make_solver(hip, N, lab, body, ppc, seed, flags=solver_flags, step_flags=other)
        torch.cuda.synchronize()
        a0 = s2.stage_times()
        t0 = time.perf_counter()
        s2.step(K)
        torch.cuda.synchronize()
        e2 = time.perf_counter() - t0
        a1 = s2.stage_times()
        result["other_coupling"] = {
            "coupling": "SDF face fractions" if other & FS_STEP_FACE_FRACTIONS else
                        "voxels at cell centres (the reference's)",
            "value": K / e2, "unit": "steps/s", "ms_per_step": 1e3 * e2 / K,
            "stage_ms": {k: round(v / K, 4) for k, v in stage_delta(a0, a1).items()},
            "note": f"same scene and seed on a fresh handle, steps 1..{K} in one fs_step call"}
        del s2
        torch.cuda.empty_cache()

    if not args.no_cpu:
        # the CPU leg: the oracle timed on a bounded sample of the workload, and
        # the oracle's FP64 PCG as the checker of the last GPU pressure system
        from fluidprjct_amd._abi import FsLib
        orc = FsLib(os.path.join(ROOT, "oracle", "_build", "libfs_oracle.so"))
        threads = int(orc.lib.oracle_num_threads()) if hasattr(orc.lib, "oracle_num_threads") else 1
        log(f"FP64 reference solve of the last step's system (oracle, {threads} threads)")
        t0 = time.perf_counter()
        if step_flags & FS_STEP_FACE_FRACTIONS:
            # the fraction-weighted system of the body's final state
            o = orc.solver(dx=1.0 / N, solver_tol=1e-12, solver_max_iter=100000, step_flags=step_flags)
            o.set_labels(lab)
            o.set_body_sdf(body["sdf"], body["origin"], body["cell_size"], body["pos"], body["mass"])
            o.set_body_state(sys_body)
            # the rotation the step used (it lags the angles by one integration)
            import ctypes
            fn = orc.lib.oracle_set_body_rotation
            fn.restype, fn.argtypes = ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
            orc.check(fn(o.h, (ctypes.c_float * 9)(*sys_body.rotation)))
            o.set_grid(FS_GRID_LABELS, sys_lab)
            o.set_grid(FS_GRID_RHS, sys_b)
            o.stage("pressure")
            pref, sref = o.get_grid(FS_GRID_P64), o.solve_stats()
        else:
            o = orc.solver()
            pref, sref = o.solve_pressure_system(sys_lab, sys_b, 1e-12)
        tref = time.perf_counter() - t0
        fl = sys_lab == 1
        dp = np.where(fl, sys_p - pref, 0.0)
        scale = sys_dt / (1000.0 * (1.0 / N))   # dt / (rho dx): the projection's pressure-to-velocity factor
        result["error_vs_fp64"] = {
            "pressure_rel_l2": float(np.linalg.norm(dp[fl]) / np.linalg.norm(pref[fl])),
            "velocity_rel_l2": float(scale * face_gradient_norm(dp, sys_lab) / u_norm) if u_norm > 0 else None,
            "reference": f"oracle FP64 MIC(0)-PCG to 1e-12 relative residual ({sref.iterations} iterations, "
                         f"{tref:.1f} s); system of GPU step {K + 1} (the step after the timed ones)",
            "velocity_note": "dt/(rho dx) |grad(p_gpu - p_fp64)| over the faces the projection updates, "
                             "relative to |u| on the faces of fluid cells after projection",
        }
        del o
        log(f"CPU baseline: {args.cpu_steps} oracle step(s)")
        o = make_solver(orc, N, lab, body, ppc, seed, step_flags=step_flags)
        t0 = time.perf_counter()
        for i in range(args.cpu_steps):
            o.step(1)
            log(f"  oracle step {i + 1} done")
        ct = time.perf_counter() - t0
        result["cpu_baseline"] = {
            "value": args.cpu_steps / ct, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"first {args.cpu_steps} step(s) of the same workload from the same seeded state (oracle "
                      f"C++ -O3; OpenMP {threads} threads in labels, RHS, projection, G2P/advection and the "
                      f"CG's SpMV and dots; P2G, the MIC(0) triangular solves and the body sums serial; FP64 "
                      f"MIC(0)-PCG to 1e-6)",
            "cpu_model": cpu_model(),
            "host_cpus": len(os.sched_getaffinity(0)),
            "stage_ms": {k: round(v / args.cpu_steps, 1) for k, v in o.stage_times().items()},
        }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
