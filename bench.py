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


def make_solver(lib, N, lab, body, ppc, seed, tol=1e-6, flags=0):
    s = lib.solver(dt=1.0 / 120, dx=1.0 / N, pic_weight=0.05, cfl=1.0, solver_tol=tol, solver_max_iter=400,
                   particles_per_cell=ppc, seed=seed, solver_flags=flags)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch the inner PCG iterations eagerly")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--learned-steps", type=int, default=5)
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
    hip = load_library()

    N, lab, body, ppc, seed, desc = build_scene(args.config)
    s = make_solver(hip, N, lab, body, ppc, seed, flags=1 if args.no_graph else 0)
    Np = s.num_particles
    s.step(args.warmup)
    torch.cuda.synchronize()
    t_stage0 = s.stage_times()
    t0 = time.perf_counter()
    s.step(args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    t_stage = stage_delta(t_stage0, s.stage_times())
    K = args.steps
    st = s.solve_stats()
    simt, dt_last, nsteps = s.clock()

    D, H, W = lab.shape
    C = W * H * D
    F = ((W + 1) * H * D + W * (H + 1) * D + W * H * (D + 1)) // 3   # faces per component (mean)
    labs = s.get_grid(7)
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
        "config": {"workload": args.config + ": " + desc, "grid": [W, H, D], "particles": Np,
                   "fluid_cells": nf, "precision": "FP32 transfers/stencil, FP64 dots + refinement"},
        "pressure_solve_ms_plain": ms["pressure"],
        "pressure_iters_last": st.iterations,
        "pressure_outer_last": st.outer,
        "pressure_rel_residual_last": st.rel_residual,
        "stage_ms": {k: round(v, 4) for k, v in ms.items()},
        "particles_per_s_transfers": Np / ((ms["p2g"] + ms["g2p_advect"]) * 1e-3),
        "dt_last": dt_last,
        "roofline": {"bound": "hbm", "kernel": KERNEL_NAMES[dom],
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(KERNEL_NAMES[dom], args.config), "alg_bytes": alg[dom],
                     "avg_launch_ms": ms[dom]},
    }
    # learned vs plain: the same workload continued with the learned
    # preconditioner (flexible PCG, multigrid fallback after 5 stalled iterations)
    if not args.no_learned:
        from fluidprjct_amd.learned import DEFAULT_PARAMS, use_learned
        use_learned(s, DEFAULT_PARAMS, 5)
        s.step(1)
        torch.cuda.synchronize()
        t0s = s.stage_times()
        fb = 0
        iters = []
        for _ in range(args.learned_steps):
            s.step(1)
            st = s.solve_stats()
            fb += st.fallback
            iters.append(st.iterations)
        torch.cuda.synchronize()
        tl = stage_delta(t0s, s.stage_times())
        result["pressure_solve_ms_learned"] = tl["pressure"] / args.learned_steps
        result["learned"] = {"steps": args.learned_steps, "fallbacks": fb, "iterations": iters,
                             "params": os.path.relpath(DEFAULT_PARAMS, ROOT),
                             "note": "same workload continued after the timed steps; flexible PCG with the "
                                     "BF16 conv net, multigrid fallback after 5 stalled iterations"}
    del s
    torch.cuda.empty_cache()

    if not args.no_cpu:
        from fluidprjct_amd._abi import FsLib
        orc = FsLib(os.path.join(ROOT, "oracle", "_build", "libfs_oracle.so"))
        o = make_solver(orc, N, lab, body, ppc, seed)
        t0 = time.perf_counter()
        o.step(args.cpu_steps)
        ct = time.perf_counter() - t0
        threads = int(orc.lib.oracle_num_threads()) if hasattr(orc.lib, "oracle_num_threads") else 1
        result["cpu_baseline"] = {
            "value": args.cpu_steps / ct, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"first {args.cpu_steps} step(s) of the same workload from the same seeded state "
                      f"(oracle C++ -O3, OpenMP {threads} threads, FP64 MIC(0)-PCG to 1e-6)",
            "stage_ms": {k: round(v / args.cpu_steps, 1) for k, v in o.stage_times().items()},
        }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
