"""Benchmark: SME level 2 + implicit Newtonian-slip source, FP64, one MI355X.

Workload (the headline configuration): 1000 x 1000 unit-square grid of right triangles
(2 000 000 cells), 16 seeded Gaussian bumps, h = 1 - b + 0.1 * circular dam, MUSCL with the
Vacuum + Barth-Jespersen (TVD) limiter, Green-Gauss gradient, Rusanov + hydrostatic
reconstruction + non-conservative fluctuations, SSP-RK2, implicit source Newton per cell,
local CFL 0.45.  One step = the full MUSCLSolver time step (2 RHS stages, constraints,
source, constraints, CFL reduction) over all cells, with the state resident in HBM.

Prints ONE JSON line with the metric, the HBM roofline of the dominant kernel (flux) from
HIP events on the solver stream, and a CPU baseline (the oracle, 1 thread, bounded sample).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "zoomy-dmplex_amd")

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, GB/s


def load_package():
    spec = importlib.util.spec_from_file_location(
        "zoomy_dmplex_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["zoomy_dmplex_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def flux_pass_bytes_per_cell(n_q, faces_per_cell):
    """Algorithmic HBM bytes of one flux pass per cell: read Q (8n), gradient (16n),
    bounds (16n), cell geometry (24), connectivity (24), face geometry (40 F);
    write the stage state (8n)  ->  48 n + 48 + 40 F."""
    return 48 * n_q + 48 + 40 * faces_per_cell


def step_bytes_per_cell(n_q, faces_per_cell, stages=2):
    """Algorithmic HBM bytes of one full step per cell:
    s (88 n + 96 + 56 F) + 8 n (Q^n in the last stage) + 16 n (source) + 8 n + 8 (CFL)."""
    return stages * (88 * n_q + 96 + 56 * faces_per_cell) + 32 * n_q + 8


def cpu_baseline(zdm, cases, grid=250, steps=20):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import oracle  # CPU checker, timed here as the reported baseline
    orc = oracle("sme2")
    mesh = zdm.Mesh.tri_grid(grid, grid)
    info = zdm.ModelInfo("sme2")
    X = cases.c2_state(mesh, info.n_q, seed=2)
    A = np.zeros((mesh.ncells, info.n_aux))
    fb = zdm.boundary_face_bc(mesh, info.boundary_tags, info.dim)
    p = info.default_params
    # slope method: time T and 2T steps, report d(wall)/d(steps)
    runs = []
    for n in (steps, 2 * steps):
        Xc, Ac = X.copy(), A.copy()
        t0 = time.perf_counter()
        orc.run(mesh, fb, p, Xc, Ac, order=2, ssp_stages=2, cfl=0.45, max_steps=n)
        runs.append(time.perf_counter() - t0)
    per_step = (runs[1] - runs[0]) / steps
    return {
        "value": mesh.ncells / per_step / 1e6,
        "unit": "Mcell-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": "oracle (C++ restatement of the reference step, 1 thread) on the same "
                  "workload at %dx%d (%d cells), slope of %d vs %d steps: %.1f ns/cell/step"
                  % (grid, grid, mesh.ncells, steps, 2 * steps, per_step / mesh.ncells * 1e9),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=1000)
    ap.add_argument("--order", default="rcm", choices=["hilbert", "rcm", "none"],
                    help="cell renumbering of the mesh service")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d refused)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    zdm = load_package()
    import zoomy_dmplex_amd.cases as cases

    n = args.grid
    mesh = zdm.Mesh.tri_grid(n, n, renumber=args.order)
    info = zdm.ModelInfo("sme2")
    X = cases.c2_state(mesh, info.n_q, seed=2)
    s = zdm.Solver("sme2", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=True,
                   gradient="green_gauss", use_limiters=True)
    s.upload(X, None)
    s.initialize()
    s.step(args.warmup)
    s.synchronize()

    s.enable_timing(True)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    flux_ms, flux_n = s.kernel_time(1)
    prep_ms, prep_n = s.kernel_time(0)
    post_ms, post_n = s.kernel_time(2)
    s.enable_timing(False)
    t_sim, dt, nsteps = s.time()
    Xg, _ = s.download()
    if not np.all(np.isfinite(Xg)):
        raise RuntimeError("non-finite state after the benchmark")

    ncells = mesh.ncells
    F = mesh.nfaces / ncells
    value = ncells * args.steps / elapsed / 1e6
    flux_bytes = flux_pass_bytes_per_cell(info.n_q, F) * ncells
    achieved = flux_bytes / (flux_ms * 1e-3) / 1e9
    out = {
        "metric": "Mcell-steps/s (SME L2, FP64)",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "SME level 2 + implicit slip source, %dx%d right-triangle grid "
                        "(%d cells), MUSCL + Barth-Jespersen, SSP-RK2, CFL 0.45" % (n, n, ncells),
            "cells": ncells,
            "cell_order": args.order,
            "n_q": info.n_q,
            "faces_per_cell": round(F, 4),
            "step_bytes_per_cell": round(step_bytes_per_cell(info.n_q, F), 1),
            "kernel_ms": {"prep": round(prep_ms, 4), "flux": round(flux_ms, 4),
                          "post": round(post_ms, 4)},
            "sim_time": t_sim,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_flux (flux + fluctuations gather + SSP stage update)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "bytes_per_cell": flux_pass_bytes_per_cell(info.n_q, F),
            "launches": flux_n,
        },
    }
    copy_gbs, _ = zdm.stream_copy_gbs("sme2", 1 << 30, 20)
    out["roofline"]["stream_copy_gbs"] = round(copy_gbs, 1)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(zdm, cases)
    s.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
