"""Benchmark: SME level 2 + implicit Newtonian-slip source, FP64, one MI355X.

Workload (the headline configuration): 1000 x 1000 unit-square grid of right triangles
(2 000 000 cells), 16 seeded Gaussian bumps, h = 1 - b + 0.1 * circular dam, MUSCL with the
Vacuum + Barth-Jespersen (TVD) limiter, Green-Gauss gradient, Rusanov + hydrostatic
reconstruction + non-conservative fluctuations, SSP-RK2, implicit source Newton per cell,
local CFL 0.45.  One step = the full MUSCLSolver time step (2 RHS stages, constraints,
source, constraints, CFL reduction) over all cells, with the state resident in HBM.

Prints ONE JSON line with the metric, the HBM roofline of the dominant kernel (flux) from
HIP events on the solver stream, and a CPU baseline (the oracle, 1 thread, bounded sample).

--workload c3 measures the VAM Chorin pressure split instead: 707 x 707 grid jittered by
+-15 % (999 698 cells), solitary wave of amplitude 0.2 over a seeded bump field, each step =
predictor + GMRES(30) / point-block Jacobi to rtol 1e-10 + corrector.
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
TIMING_STRIDE = 10      # C2: per-kernel HIP events on every 10th timed step


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


def flux_launch_bytes_per_cell(n_q, faces_per_cell, stages=2):
    """Mean algorithmic HBM bytes per cell of one k_flux launch of a step: s - 1 stage launches
    of a flux pass and the final one, which also reads Q^n for the SSP combination (the step's
    "8 n (Q^n in the last stage)" term): 48 n + 48 + 40 F + 8 n / s."""
    return flux_pass_bytes_per_cell(n_q, faces_per_cell) + 8.0 * n_q / stages


def final_post_bytes_per_cell(n_q, faces_per_cell):
    """Algorithmic HBM bytes per cell of the final stage fused with the post-step (k_flux_post):
    a flux pass, Q^n for the SSP combination (8 n), the source's read and write of the state
    (16 n) and the CFL pass (8 n + 8), the step's terms of those passes; the kernel itself moves
    less, since the new state stays in registers between them."""
    return flux_pass_bytes_per_cell(n_q, faces_per_cell) + 32 * n_q + 8


def step_bytes_per_cell(n_q, faces_per_cell, stages=2):
    """Algorithmic HBM bytes of one full step per cell:
    s (88 n + 96 + 56 F) + 8 n (Q^n in the last stage) + 16 n (source) + 8 n + 8 (CFL)."""
    return stages * (88 * n_q + 96 + 56 * faces_per_cell) + 32 * n_q + 8


def pmc_traffic(kernel, cells):
    """Bytes per launch of `kernel` past the L2 (an upper bound on its HBM bytes; k_flux: the mean of
    its stage and final launches) from
    the committed FETCH_SIZE / WRITE_SIZE profile of the same workload
    (profiles/k_flux_c2_2M_pmc.md, copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "k_flux_c2_2M_pmc.md")
    try:
        for line in open(path):
            if line.startswith("bench: "):
                d = json.loads(line[len("bench: "):])
                if d["kernel"] == kernel and d["cells"] == cells:
                    return d["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        pass
    return None


_CPU_CHILD = r"""
import sys, time, json
sys.path.insert(0, sys.argv[1])
from conftest import oracle, zdm, cases
import numpy as np
workload, arg, steps = sys.argv[2], int(sys.argv[3]), int(sys.argv[4])
extra = json.loads(sys.argv[5]) if len(sys.argv) > 5 else {}
if workload == "c3":
    orc = oracle("vam")
    mesh = zdm.Mesh.tri_grid(arg, arg, jitter=0.15, seed=3)
    X = cases.c3_state(mesh)
    A = np.zeros((mesh.ncells, orc.n_aux))
    fb = zdm.boundary_face_bc(mesh, ["default"], 2)
    runs, its = [], []
    for n in (steps, 2 * steps):
        Xc, Ac, Pc = X.copy(), A.copy(), np.zeros((mesh.ncells, 2))
        t0 = time.perf_counter()
        _, _, it, reason = orc.chorin_run(mesh, fb, orc.default_params, Xc, Ac, Pc, order=1, max_steps=n, **extra)
        runs.append(time.perf_counter() - t0)
        its.append(it)
    print(json.dumps({"cells": mesh.ncells, "per_step": (runs[1] - runs[0]) / steps,
                      "its_per_step": (its[1] - its[0]) / steps}))
    sys.exit(0)
if workload == "c2":
    model = "sme2"
    mesh = zdm.Mesh.tri_grid(arg, arg)
    info = zdm.ModelInfo(model)
    X = cases.c2_state(mesh, info.n_q, seed=2)
elif workload == "c1":
    model = "sme0"
    mesh = zdm.Mesh.tri_grid(arg, arg)
    info = zdm.ModelInfo(model)
    X = cases.c1_state(mesh, 4, seed=1)
else:
    model = "sme1"
    n, mesh_seed, state = cases.ensemble_member(arg)
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=mesh_seed)
    info = zdm.ModelInfo(model)
    X = state(mesh)
orc = oracle(model)
A = np.zeros((mesh.ncells, info.n_aux))
fb = zdm.boundary_face_bc(mesh, info.boundary_tags, info.dim)
runs = []
for n in (steps, 2 * steps):
    Xc, Ac = X.copy(), A.copy()
    t0 = time.perf_counter()
    orc.run(mesh, fb, info.default_params, Xc, Ac, order=2, ssp_stages=2, cfl=0.45, max_steps=n)
    runs.append(time.perf_counter() - t0)
print(json.dumps({"cells": mesh.ncells, "per_step": (runs[1] - runs[0]) / steps}))
"""

# The measuring host gives each GPU a share of 16 cores, whatever os.cpu_count() reports there
# (it counts the whole machine): the CPU baselines run at most that many processes.
CPU_SHARE_PER_GPU = 16


def host_cpu():
    """os.cpu_count() and the CPU model of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"os_cpu_count": os.cpu_count(), "cpu_model": model}


def _oracle_procs(args_list, steps, extra=None):
    """One oracle process per entry of args_list (workload, arg), all at once."""
    import subprocess
    env = dict(os.environ, ZDM_NO_TORCH="1", OMP_NUM_THREADS="1")
    tail = [json.dumps(extra)] if extra else []
    ps = [subprocess.Popen([sys.executable, "-c", _CPU_CHILD, os.path.join(ROOT, "tests"), w, str(a), str(steps)] + tail,
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env, text=True)
          for w, a in args_list]
    res = []
    for p in ps:
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("CPU baseline process failed: " + err[-2000:])
        res.append(json.loads(out.strip().splitlines()[-1]))
    return res


def cpu_baseline(zdm, cases, grid=250, steps=20, procs=None, workload="c2"):
    """The oracle (C++ restatement of the reference step, one thread each) run as `procs`
    concurrent processes on the host's cores (at most the 16-core share of one GPU), each on
    the same bounded sample; the aggregate throughput is reported (slope of T vs 2T steps
    removes start-up)."""
    procs = procs or max(1, min(CPU_SHARE_PER_GPU, os.cpu_count() or 1))
    res = _oracle_procs([(workload, grid)] * procs, steps)
    cells = res[0]["cells"]
    per_step = sum(r["per_step"] for r in res) / len(res)
    return dict({
        "value": procs * cells / per_step / 1e6,
        "unit": "Mcell-steps/s",
        "cores": procs,
        "kind": "port",
        "sample": "oracle (C++ restatement of the reference step, one thread per process) as %d "
                  "concurrent processes (the 16-core share of one GPU on the measuring host), each on "
                  "the same workload at %dx%d (%d cells), slope of %d vs %d steps: %.1f ns/cell/step "
                  "per process" % (procs, grid, grid, cells, steps, 2 * steps, per_step / cells * 1e9),
    }, **host_cpu())


def c4_cpu_baseline(zdm, cases, members, steps=10, procs=None):
    """The oracle on the ensemble's own members, one member per process, `procs` processes at
    once (members spread over the ensemble's index range): aggregate Mcell-steps/s."""
    procs = procs or max(1, min(CPU_SHARE_PER_GPU, os.cpu_count() or 1))
    pick = sorted(set(int(round(i * (members - 1) / max(procs - 1, 1))) for i in range(procs)))
    res = _oracle_procs([("c4", i) for i in pick], steps)
    value = sum(r["cells"] / r["per_step"] for r in res) / 1e6
    return dict({
        "value": value,
        "unit": "Mcell-steps/s",
        "cores": len(pick),
        "kind": "port",
        "sample": "oracle (C++ restatement of the reference step, one thread per process) on %d of the "
                  "%d members (%d..%d cells), one per process, all at once (the 16-core share of one GPU), "
                  "slope of %d vs %d steps, throughputs summed: the CPU's ensemble rate when each core "
                  "runs its own member" % (len(pick), members, min(r["cells"] for r in res),
                                           max(r["cells"] for r in res), steps, 2 * steps),
    }, **host_cpu())


def c5_cpu_baseline(zdm, cases, grid=120, steps=10):
    """The C5 problem on the CPU: the oracle's forward loss and its tangent-linear run (one
    direction) on a bounded sample, one thread.  value = tangent-linear time / forward time,
    the CPU path's cost of one directional derivative (a full gradient over bed and Manning n
    would take 2 x cells of them)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from conftest import TangentOracle, oracle  # CPU checker, timed here as the reported baseline
    mesh = zdm.Mesh.tri_grid(grid, grid, jitter=0.15, seed=5)
    _, _, Xg, Ag, gauges = cases.c5_case(mesh)
    info = zdm.ModelInfo("sme1m")
    orc, tl = oracle("sme1m"), TangentOracle("sme1m")
    fb = zdm.boundary_face_bc(mesh, info.boundary_tags, info.dim)
    p = info.default_params
    dt = 0.5 * orc.compute_dt(mesh, p, Xg.copy(), Ag.copy(), 0.45)
    db, dn = c5_directions(mesh.ncells, 1)[0]
    dX, dA = np.zeros_like(Xg), np.zeros_like(Ag)
    dX[:, 0], dA[:, 1] = db, dn
    tf, tt = [], []
    for n in (steps, 2 * steps):
        dts = np.full(n, dt)
        obs = np.zeros((n // 5, len(gauges)))
        t0 = time.perf_counter()
        orc.loss(mesh, fb, p, Xg, Ag, dts, gauges, 5, obs)
        tf.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        tl.loss(mesh, fb, p, Xg, dX, Ag, dA, dts, gauges, 5, obs)
        tt.append(time.perf_counter() - t0)
    fwd, tan = (tf[1] - tf[0]) / steps, (tt[1] - tt[0]) / steps
    return dict({
        "value": round(tan / fwd, 3),
        "unit": "x forward time (one directional derivative)",
        "cores": 1,
        "kind": "port",
        "forward_mcell_steps_per_s": round(mesh.ncells / fwd / 1e6, 3),
        "sample": "oracle forward loss and its tangent-linear run (dual numbers, one direction in bed "
                  "and Manning n, exact limiter), %dx%d jittered grid (%d cells), fixed dt, slope of %d vs "
                  "%d steps, 1 thread; the CPU path has no adjoint: a full gradient takes one "
                  "tangent-linear run per parameter (%d here)"
                  % (grid, grid, mesh.ncells, steps, 2 * steps, 2 * mesh.ncells),
    }, **host_cpu())


def c3_cpu_baseline(zdm, cases, grid=60, steps=1, krylov=None, procs=None):
    """The oracle's ChorinVAMSolver (sequential GMRES, one thread each) as `procs` concurrent
    processes on the same bounded sample, like cpu_baseline; aggregate Mcell-steps/s."""
    procs = procs or max(1, min(CPU_SHARE_PER_GPU, os.cpu_count() or 1))
    res = _oracle_procs([("c3", grid)] * procs, steps, extra=krylov)
    cells = res[0]["cells"]
    per_step = sum(r["per_step"] for r in res) / len(res)
    its = res[0]["its_per_step"]
    return dict({
        "value": procs * cells / per_step / 1e6,
        "unit": "Mcell-steps/s",
        "cores": procs,
        "kind": "port",
        "mcell_its_per_s": round(procs * cells * its / per_step / 1e6, 3),
        "sample": "oracle (C++ restatement of ChorinVAMSolver, sequential GMRES, one thread per process) as "
                  "%d concurrent processes (the 16-core share of one GPU), each on a %dx%d jittered grid "
                  "(%d cells), slope of %d vs %d steps, %.1f GMRES its/step"
                  % (procs, grid, grid, cells, steps, 2 * steps, its),
    }, **host_cpu())


def run_c1(zdm, cases, args):
    """C1: SME level 0 dam break over a bump, 60 x 60 grid (7 200 cells), SSP-RK2, CFL 0.45,
    200 steps, replayed as a captured step graph (a grid this small is launch-bound)."""
    n = args.grid if args.grid != 1000 else 60
    mesh = zdm.Mesh.tri_grid(n, n, renumber=args.order)
    X = cases.c1_state(mesh, 4, seed=1)
    s = zdm.Solver("sme0", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9)
    s.upload(X, None)
    s.initialize()
    s.step_graph(args.warmup)
    s.synchronize()
    s.upload(X, None)
    s.initialize()
    s.synchronize()
    t0 = time.perf_counter()
    s.step_graph(args.steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    Xg, _ = s.download()
    s.close()
    if not np.all(np.isfinite(Xg)):
        raise RuntimeError("non-finite C1 state")
    out = {
        "metric": "Mcell-steps/s (SME L0, FP64)",
        "value": round(mesh.ncells * args.steps / dt / 1e6, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "SME level 0 dam break over a Gaussian bump, %dx%d right-triangle grid (%d cells), "
                               "MUSCL + TVD limiter, SSP-RK2, CFL 0.45, graph-replayed steps" % (n, n, mesh.ncells),
                   "cells": mesh.ncells},
        "roofline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(zdm, cases, grid=n, steps=100, workload="c1")
    print(json.dumps(out))


def run_c3(zdm, cases, args):
    n = args.grid if args.grid != 1000 else 707
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=3, renumber=args.order)
    X = cases.c3_state(mesh)
    krylov = dict(restart=args.restart, max_it=args.max_it, rtol=1e-10, atol=1e-12, op_mode=args.op_mode)
    s = zdm.Solver("vam", mesh, order=1, cfl=0.45, t_end=1e9, source=True)
    s.chorin_setup(**krylov)
    s.upload(X, None)
    s.initialize()
    steps_w, its_w, _ = s.chorin_step(args.warmup)
    s.synchronize()
    t0 = time.perf_counter()
    steps, its, reason = 0, 0, 0
    step_its, step_rel = [], []
    for _ in range(args.steps):   # one step per call: the per-step solve statistics
        n1, it1, reason = s.chorin_step(1)
        if n1 == 0:
            break
        steps += n1
        its += it1
        _, _, rn, rn0 = s.chorin_stats()
        step_its.append(it1)
        step_rel.append(rn / rn0 if rn0 > 0 else 0.0)
        print("c3 step %d: %d GMRES iterations, reason %d, relative residual %.3e"
              % (steps, it1, reason, step_rel[-1]), file=sys.stderr, flush=True)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    t_sim, dt, nsteps = s.time()
    Xg, _ = s.download()
    if not np.all(np.isfinite(Xg)):
        raise RuntimeError("non-finite state after the benchmark")
    ncells = mesh.ncells
    out = {
        "metric": "Mcell-steps/s (VAM Chorin, FP64)",
        "value": round(ncells * steps / elapsed / 1e6, 4),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": steps_w,
        "ms_per_step": round(elapsed / max(steps, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "VAM level 1 Chorin split, %dx%d grid jittered +-15%% (%d cells), "
                        "solitary wave a=0.2 over bumps <= 0.1, GMRES(%d) + PBJ rtol 1e-10"
                        % (n, n, ncells, args.restart),
            "cells": ncells,
            "op_mode": ["matrix-free residual", "assembled Pmat"][args.op_mode],
            "gmres_its_per_step": round(its / max(steps, 1), 2),
            "gmres_its_each_step": step_its,
            "last_reason": reason,
            "converged": reason > 0,
            "rel_residual_each_step": [float("%.3e" % r) for r in step_rel],
            "max_it": args.max_it,
            "ms_per_gmres_it": round(elapsed * 1e3 / max(its, 1), 4),
            "mcell_its_per_s": round(ncells * its / elapsed / 1e6, 3),
            "sim_time": t_sim,
        },
        "roofline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = c3_cpu_baseline(zdm, cases, krylov=dict(krylov, op_mode=args.op_mode))
    s.close()
    print(json.dumps(out))


def run_c3adj(zdm, cases, args):
    """Calibration gradient through the Chorin split (C3 mesh and solver): recording forward +
    reverse sweep (re-run of each step, transposed pressure solve, transposed kernels) against
    the plain forward run of the same steps."""
    n = args.grid if args.grid != 1000 else 707
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=3, renumber=args.order)
    Xt = cases.c3_state(mesh)
    Xg = cases.c3_state(mesh, amp=0.18)
    krylov = dict(restart=30, max_it=args.max_it, rtol=1e-10, atol=1e-12, op_mode=args.op_mode)
    s = zdm.Solver("vam", mesh, order=1, cfl=0.45, t_end=1e9, source=True)
    s.chorin_setup(**krylov)
    T, obs_every = args.steps, 5
    rng = np.random.default_rng(3)
    gauges = rng.choice(mesh.ncells, 32, replace=False).astype(np.int32)
    # synthetic observations: h at the gauges every 5 steps of the hidden-truth run
    s.adjoint_setup(T, gauges=gauges, obs_every=obs_every, chorin=True)
    s.adjoint_forward(Xt, None)
    obs = s.adjoint_residuals().copy()
    s.adjoint_setup(T, gauges=gauges, obs_every=obs_every, obs=obs,
                    checkpoint_every=0 if args.checkpoint_every is None else args.checkpoint_every,
                    chorin=True)
    ck, _ = s.adjoint_layout()
    s.upload(Xg, None)
    s.initialize()
    s.chorin_step(min(args.warmup, 3))
    s.upload(Xg, None)
    s.initialize()
    s.synchronize()
    t0 = time.perf_counter()
    steps, its, _ = s.chorin_step(T)
    s.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    loss = s.adjoint_forward(Xg, None)
    t_fwd = time.perf_counter() - t0
    its_t0 = s.adjoint_krylov_its()
    t0 = time.perf_counter()
    gX, _ = s.adjoint_gradient()
    t_rev = time.perf_counter() - t0
    its_t = s.adjoint_krylov_its() - its_t0
    if not np.all(np.isfinite(gX)):
        raise RuntimeError("non-finite gradient")
    ncells = mesh.ncells
    out = {
        "metric": "adjoint/forward ratio (VAM Chorin calibration, FP64)",
        "value": round((t_fwd + t_rev) / t_plain, 3),
        "unit": "x forward time",
        "n_gpus": 1,
        "steps": T,
        "warmup": min(args.warmup, 3),
        "ms_per_step": round(t_plain / T * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "VAM level 1 Chorin split, %dx%d grid jittered +-15%% (%d cells), %d steps, "
                        "32 gauges every %d steps, gradient wrt the initial state; GMRES(30) + PBJ "
                        "rtol 1e-10 forward and transposed" % (n, n, ncells, T, obs_every),
            "cells": ncells,
            "forward_s": round(t_plain, 4),
            "forward_recording_s": round(t_fwd, 4),
            "reverse_s": round(t_rev, 4),
            "gmres_its_forward": its,
            "gmres_its_transposed": its_t,
            "loss": loss,
            "grad_norm_bed": float(np.linalg.norm(gX[:, 0])),
            "grad_norm_h": float(np.linalg.norm(gX[:, 1])),
            "checkpoint_every": ck,
        },
        "roofline": None,
    }
    s.close()
    print(json.dumps(out))


def run_c2_time_integration(zdm, cases, args):
    """The C2 workload under the IMEX (ARK2) or fully implicit (BDF2, Newton-Krylov) strategy:
    throughput, the kernels' share, and for the implicit step the Newton / GMRES counts."""
    n = args.grid
    mesh = zdm.Mesh.tri_grid(n, n, renumber=args.order)
    info = zdm.ModelInfo("sme2")
    X = cases.c2_state(mesh, info.n_q, seed=2)
    ti = args.time_integration
    s = zdm.Solver("sme2", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=True,
                   gradient="green_gauss", use_limiters=args.limiter, time_integration=ti)
    s.upload(X, None)
    s.initialize()
    s.step(args.warmup)
    s.synchronize()
    n0 = s.implicit_stats() if ti == "implicit" else None
    t0 = time.perf_counter()
    s.step(args.steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    Xg, _ = s.download()
    if not np.all(np.isfinite(Xg)):
        raise RuntimeError("non-finite state after the benchmark")
    ncells = mesh.ncells
    cfg = {
        "workload": "SME level 2 + implicit slip source, %dx%d right-triangle grid (%d cells), MUSCL "
                    "(limiter %s), time_integration %s, CFL 0.45" % (n, n, ncells, "on" if args.limiter else "off", ti),
        "cells": ncells,
        "time_integration": ti,
    }
    if ti == "implicit":
        nw, ksp, last, reason = s.implicit_stats()
        cfg["newton_its_per_step"] = round((nw - n0[0]) / args.steps, 2)
        cfg["gmres_its_per_step"] = round((ksp - n0[1]) / args.steps, 2)
        cfg["last_snes_reason"] = reason
    out = {
        "metric": "Mcell-steps/s (SME L2, FP64)",
        "value": round(ncells * args.steps / elapsed / 1e6, 3),
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
        "config": cfg,
        "roofline": None,
    }
    s.close()
    print(json.dumps(out))


def c4_members(zdm, cases, count):
    meshes, states = [], []
    for i in range(count):
        n, mesh_seed, state = cases.ensemble_member(i)
        m = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=mesh_seed)
        meshes.append(m)
        states.append(state(m))
    return meshes, states


def run_c4(zdm, cases, args):
    """Batched ensemble vs the same runs one at a time (both replay a graph-captured step)."""
    count = args.members
    meshes, states = c4_members(zdm, cases, count)
    kw = dict(order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=True, use_limiters=True)
    e = zdm.Ensemble("sme1", meshes, **kw)
    ncells = e.mesh.ncells
    e.upload(e.join(states))
    e.initialize()
    bs, cpt, tune = e.autotune(3)          # launch shape of the batched step, by measurement
    e.step_graph(args.warmup)
    e.synchronize()
    t0 = time.perf_counter()
    e.step_graph(args.steps)
    e.synchronize()
    t_batch = time.perf_counter() - t0
    Xe, _ = e.download()
    if not np.all(np.isfinite(Xe)):
        raise RuntimeError("non-finite ensemble state")
    dts = [e.member_time(i)[1] for i in range(count)]
    e.close()
    # the same runs one at a time: replaying a captured step (default launch shape, and each
    # member autotuned), and with plain launches
    t_serial = t_eager = t_tuned = 0.0
    shapes = {}
    for m, X in zip(meshes, states):
        s = zdm.Solver("sme1", m, **kw)
        s.upload(X)
        s.initialize()
        s.step_graph(args.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.step_graph(args.steps)
        s.synchronize()
        t_serial += time.perf_counter() - t0
        s.upload(X)
        s.initialize()
        mb, mc, _ = s.autotune(3)
        shapes[(mb, mc)] = shapes.get((mb, mc), 0) + 1
        s.step_graph(args.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.step_graph(args.steps)
        s.synchronize()
        t_tuned += time.perf_counter() - t0
        s.set_launch(256, 1)
        s.upload(X)
        s.initialize()
        s.step(args.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(args.steps)
        s.synchronize()
        t_eager += time.perf_counter() - t0
        s.close()
    F = e.mesh.nfaces / ncells
    bpc = step_bytes_per_cell(6, F)
    value = ncells * args.steps / t_batch / 1e6
    out = {
        "metric": "batched vs serial ensemble (SME L1, FP64)",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batch / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%d SME level 1 members, jittered split grids of %d..%d cells (%d in all), "
                        "MUSCL + Barth-Jespersen, SSP-RK2, CFL 0.45, own dt per member"
                        % (count, min(m.ncells for m in meshes), max(m.ncells for m in meshes), ncells),
            "members": count,
            "cells": ncells,
            "serial_mcell_steps_per_s": round(ncells * args.steps / t_serial / 1e6, 3),
            "batched_s": round(t_batch, 4),
            "serial_s": round(t_serial, 4),
            "speedup_batched_vs_serial": round(t_serial / t_batch, 3),
            "serial_eager_s": round(t_eager, 4),
            "speedup_batched_vs_serial_eager": round(t_eager / t_batch, 3),
            "serial_autotuned_s": round(t_tuned, 4),
            "speedup_batched_vs_serial_autotuned": round(t_tuned / t_batch, 3),
            "batched_launch": {"block_size": bs, "cells_per_thread": cpt,
                               "autotune_ms_per_step": {"%dx%d" % k: round(v, 4) for k, v in tune.items()}},
            "serial_autotuned_shapes": {"%dx%d" % k: v for k, v in sorted(shapes.items())},
            "dt_range": [min(dts), max(dts)],
            "step_bytes_per_cell": round(bpc, 1),
            "step_gbs": round(bpc * ncells * args.steps / t_batch / 1e9, 1),
        },
        "roofline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = c4_cpu_baseline(zdm, cases, count)
    print(json.dumps(out))


def c5_problem(zdm, cases, args):
    n = args.grid if args.grid != 1000 else 354
    mesh = zdm.Mesh.tri_grid(n, n, jitter=0.15, seed=5, renumber=args.order)
    Xt, At, Xg, Ag, gauges = cases.c5_case(mesh)
    s = zdm.Solver("sme1m", mesh, order=2, ssp_stages=2, cfl=0.45, t_end=1e9, source=True,
                   use_limiters=True)
    return n, mesh, s, Xt, At, Xg, Ag, gauges


def c5_observations(s, T, gauges, obs_every, Xt, At):
    """h at the gauges every obs_every steps of the hidden-truth run."""
    s.adjoint_setup(T, gauges=gauges, obs_every=obs_every)
    s.adjoint_forward(Xt, At)
    return s.adjoint_residuals().copy()


def c5_default_layout(T):
    """The configured layout: the binomial schedule on ceil(sqrt(T)) state slots."""
    return -2 - int(np.ceil(np.sqrt(T)))


def _device_arrays(*arrays):
    """float64 device copies (torch) of host arrays, or None without a GPU torch."""
    try:
        import torch
        if not torch.cuda.is_available():
            return None
        out = [torch.tensor(a, dtype=torch.float64, device="cuda") for a in arrays]
        torch.cuda.synchronize()
        return out
    except ImportError:
        return None


def c5_measure(s, T, gauges, obs_every, obs, X, A, ck, exact, stretch=None):
    """Recording forward + reverse sweep, once to warm up and once timed, with the inputs and
    the gradient in device memory (zdm_adjoint_forward_device / _gradient_device) when torch
    has the GPU, else through host arrays."""
    s.adjoint_setup(T, gauges=gauges, obs_every=obs_every, obs=obs, checkpoint_every=ck,
                    exact_limiter=exact, stretch_tape=stretch)
    dev = _device_arrays(X, A, np.zeros_like(X), np.zeros_like(A))
    if dev is not None:
        Xd, Ad, gXd, gAd = dev
        fwd = lambda: s.adjoint_forward_device(Xd.data_ptr(), Ad.data_ptr())
        rev = lambda: s.adjoint_gradient_device(gXd.data_ptr(), gAd.data_ptr())
    else:
        fwd = lambda: s.adjoint_forward(X, A)
        rev = lambda: s.adjoint_gradient()
    fwd()
    rev()
    s.synchronize()
    t0 = time.perf_counter()
    loss = fwd()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    rev()
    s.synchronize()
    t_rev = time.perf_counter() - t0
    if dev is not None:
        gX, gA = gXd.cpu().numpy(), gAd.cpu().numpy()
    else:
        gX, gA = s.adjoint_gradient()
    if not (np.all(np.isfinite(gX)) and np.all(np.isfinite(gA))):
        raise RuntimeError("non-finite gradient (checkpoint_every %d, exact %s)" % (ck, exact))
    lay = s.adjoint_layout()
    st = s.adjoint_stretch_tape()
    return {"t_fwd": t_fwd, "t_rev": t_rev, "loss": loss, "gX": gX, "gA": gA, "layout": lay, "stretch": st,
            "inputs": "device" if dev is not None else "host"}


def c5_layout_name(m):
    ck, tape = m["layout"]
    L, lv, nb = m["stretch"]
    if ck < -2:
        name = "binomial schedule on %d state slots" % (-2 - ck)
    elif tape:
        return ["", "tape of stage states", "tape of stage and face states"][tape]
    elif ck == 1:
        name = "every state"
    else:
        name = "checkpoint every %d steps" % ck
    if L:
        name += ", stretch tape of %d steps (%s)%s" % (L, "stage and face states" if lv == 2 else "stage states",
                                                       ", two buffers, re-runs on a second stream" if nb == 2 else "")
    return name


def run_c5(zdm, cases, args):
    """Calibration gradient: forward run + adjoint sweep vs the plain forward run, on the
    binomial schedule over ceil(sqrt(T)) state slots by default."""
    T = args.steps
    obs_every = 5
    n, mesh, s, Xt, At, Xg, Ag, gauges = c5_problem(zdm, cases, args)
    obs = c5_observations(s, T, gauges, obs_every, Xt, At)
    ck = c5_default_layout(T) if args.checkpoint_every is None else args.checkpoint_every
    exact = bool(args.exact_limiter)
    # plain forward of the first guess (the same steps, no recording), at the launch shape the
    # autotuner picks for it (the recording forward and the re-runs use the same shape)
    s.upload(Xg, Ag)
    s.initialize()
    bs, cpt, _ = s.autotune(3)
    s.step(min(args.warmup, T))
    s.synchronize()
    s.upload(Xg, Ag)
    s.initialize()
    s.synchronize()
    t0 = time.perf_counter()
    s.step(T)
    s.synchronize()
    t_plain = time.perf_counter() - t0
    print("c5: plain forward %.4f s" % t_plain, file=sys.stderr, flush=True)
    head = c5_measure(s, T, gauges, obs_every, obs, Xg, Ag, ck, exact)
    print("c5: %s: %.4f + %.4f s" % (c5_layout_name(head), head["t_fwd"], head["t_rev"]), file=sys.stderr, flush=True)
    breakdown = {}
    if not args.no_breakdown:
        # the same gradient from other layouts: the full tape's reverse sweep runs only the
        # transposed kernels, so the difference is the stretches' recording re-runs
        for key, ck2, ex2, st2 in (("full_tape", -2, exact, None),
                                   ("one_tape_buffer", ck, exact, (-(-T // max(1, -2 - ck)) if ck < -2 else ck, 1)),
                                   ("binomial_without_stretch_tape", ck, exact, 0),
                                   ("other_limiter_mode", ck, not exact, None)):
            m = c5_measure(s, T, gauges, obs_every, obs, Xg, Ag, ck2, ex2, st2)
            same = bool(np.array_equal(m["gX"], head["gX"]) and np.array_equal(m["gA"], head["gA"]))
            breakdown[key] = {
                "layout": c5_layout_name(m),
                "exact_limiter": ex2,
                "forward_recording_s": round(m["t_fwd"], 4),
                "reverse_s": round(m["t_rev"], 4),
                "ratio": round((m["t_fwd"] + m["t_rev"]) / t_plain, 3),
                "gradient_bitwise_equal_to_headline": same,
            }
            print("c5: %s: %.4f + %.4f s" % (key, m["t_fwd"], m["t_rev"]), file=sys.stderr, flush=True)
        ft, ob = breakdown["full_tape"], breakdown["one_tape_buffer"]
        breakdown["split_of_headline"] = {
            "plain_forward_s": round(t_plain, 4),
            "forward_recording_s": round(head["t_fwd"], 4),
            "transposed_kernels_s": ft["reverse_s"],
            "stretch_re_runs_s_one_buffer": round(ob["reverse_s"] - ft["reverse_s"], 4),
            "reverse_saved_by_overlap_s": round(ob["reverse_s"] - head["t_rev"], 4),
        }
    gX, gA = head["gX"], head["gA"]
    ncells = mesh.ncells
    ratio = (head["t_fwd"] + head["t_rev"]) / t_plain
    out = {
        "metric": "adjoint/forward ratio (SME L1 + Manning calibration, FP64)",
        "value": round(ratio, 3),
        "unit": "x forward time",
        "n_gpus": 1,
        "steps": T,
        "warmup": args.warmup,
        "ms_per_step": round(t_plain / T * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "SME level 1 + Manning, %dx%d grid jittered +-15%% (%d cells), %d SSP-RK2 "
                        "steps, 32 gauges every %d steps, gradient wrt bed and Manning n"
                        % (n, n, ncells, T, obs_every),
            "cells": ncells,
            "limiter": "exact (alpha differentiated through its active branch)" if exact else
                       "frozen (alpha held at its forward value)",
            "forward_s": round(t_plain, 4),
            "forward_recording_s": round(head["t_fwd"], 4),
            "reverse_s": round(head["t_rev"], 4),
            "reverse_over_forward": round(head["t_rev"] / t_plain, 3),
            "forward_mcell_steps_per_s": round(ncells * T / t_plain / 1e6, 2),
            "loss": head["loss"],
            "grad_norm_bed": float(np.linalg.norm(gX[:, 0])),
            "grad_norm_manning": float(np.linalg.norm(gA[:, 1])),
            "checkpoint_every": head["layout"][0],
            "layout": c5_layout_name(head),
            "launch": {"block_size": bs, "cells_per_thread": cpt},
            "inputs_and_gradient": head["inputs"] + " memory",
            "breakdown": breakdown,
        },
        "roofline": None,
    }
    s.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = c5_cpu_baseline(zdm, cases)
    print(json.dumps(out))


def c5_directions(ncells, k, seed=11, kind="white", mesh=None):
    """k seeded directions in (bed, Manning n).  white: 0.01 N(0, 1) per cell on the bed and
    0.005 N(0, 1) on n; smooth: the same amplitudes on smooth random fields (sums of random
    Fourier modes, as the unknown fields themselves)."""
    out = []
    if kind == "smooth":
        from zoomy_dmplex_amd.cases import SplitMix64, smooth_field
        rng = SplitMix64(seed)
        xy = mesh.cell_centroid[:, :2]
        for _ in range(k):
            out.append((0.01 * smooth_field(xy, rng), 0.005 * smooth_field(xy, rng)))
        return out
    rng = np.random.default_rng(seed)
    for _ in range(k):
        out.append((0.01 * rng.standard_normal(ncells), 0.005 * rng.standard_normal(ncells)))
    return out


def run_c5_fd(zdm, cases, args):
    """Central differences of the device's own frozen-dt forward run against
    <gradient, direction> on --directions random directions (white and smooth), for each FD
    step in --fd-steps (the configuration: 1e-6), at each horizon in --horizons (steps), in
    both limiter modes, with the configured layout.  --frozen-map adds, for the frozen limiter,
    the differences of the map that gradient differentiates: the forward with every face's
    limiter value held at the unperturbed run's (zdm_adjoint_forward_frozen_alpha)."""
    horizons = [int(v) for v in args.horizons.split(",")]
    epss = [float(v) for v in args.fd_steps.split(",")]
    obs_every = 5
    n, mesh, s, Xt, At, Xg, Ag, gauges = c5_problem(zdm, cases, args)
    dirsets = {kind: c5_directions(mesh.ncells, args.directions, kind=kind, mesh=mesh)
               for kind in ("smooth", "white")}
    rows = []
    for T in horizons:
        oe = min(obs_every, T)
        obs = c5_observations(s, T, gauges, oe, Xt, At)
        modes = [(True, "plain"), (False, "plain")] + ([(False, "frozen")] if args.frozen_map else [])
        for exact, fmap in modes:
            s.adjoint_setup(T, gauges=gauges, obs_every=oe, obs=obs, checkpoint_every=c5_default_layout(T),
                            exact_limiter=exact)
            loss = s.adjoint_forward(Xg, Ag)
            dts = s.adjoint_dts().copy()
            gX, gA = s.adjoint_gradient()
            if fmap == "frozen":
                s.adjoint_forward_frozen_alpha(Xg, Ag, dts=dts, record=True)
                fwd = lambda X_, A_: s.adjoint_forward_frozen_alpha(X_, A_, dts=dts)
            else:
                fwd = lambda X_, A_: s.adjoint_forward(X_, A_, dts=dts)
            for kind, dirs in dirsets.items():
                for eps in epss:
                    errs = []
                    for db, dn in dirs:
                        Xp, Xm, Ap, Am = Xg.copy(), Xg.copy(), Ag.copy(), Ag.copy()
                        Xp[:, 0] += eps * db
                        Xm[:, 0] -= eps * db
                        Ap[:, 1] += eps * dn
                        Am[:, 1] -= eps * dn
                        fd = (fwd(Xp, Ap) - fwd(Xm, Am)) / (2 * eps)
                        ad = float(np.dot(gX[:, 0], db) + np.dot(gA[:, 1], dn))
                        errs.append(abs(ad - fd) / max(abs(fd), 1e-300))
                    errs = np.array(errs)
                    row = {"steps": T, "exact_limiter": exact, "fd_map": fmap, "directions": kind, "fd_step": eps,
                           "loss": loss,
                           "grad_norm": float(np.sqrt(np.sum(gX[:, 0] ** 2) + np.sum(gA[:, 1] ** 2))),
                           "rel_err_median": float(np.median(errs)), "rel_err_max": float(np.max(errs)),
                           "within_1e-6": int(np.sum(errs <= 1e-6)), "count": len(errs)}
                    rows.append(row)
                    print("c5fd:", json.dumps(row), file=sys.stderr, flush=True)
    s.close()
    print(json.dumps({
        "metric": "adjoint vs central differences (SME L1 + Manning calibration, FP64)",
        "value": None,
        "unit": "relative error per row",
        "n_gpus": 1, "steps": horizons[-1] if horizons else 0, "warmup": 0, "higher_is_better": False,
        "dtype": "f64", "data": "synthetic", "vs_baseline": None,
        "config": {"workload": "C5 calibration problem on %d cells, central differences on %d seeded "
                               "directions per kind in (bed, Manning n), dt frozen at the forward run"
                               % (mesh.ncells, args.directions),
                   "rows": rows},
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "c5fd", "c3adj"],
                    help="c1: SME level 0 on 60x60 (the CPU case); c2: SME level 2 transport + source (headline); "
                         "c3: VAM Chorin split; "
                         "c4: batched vs serial ensemble; c5: adjoint calibration gradient; "
                         "c5fd: its central-difference check; c3adj: adjoint through the Chorin split")
    ap.add_argument("--members", type=int, default=256, help="c4 ensemble size")
    ap.add_argument("--op-mode", type=int, default=1, choices=[0, 1],
                    help="c3 pressure operator: 1 the device-assembled Pmat (sliced ELL SpMV, "
                         "default), 0 the matrix-free residual")
    ap.add_argument("--max-it", type=int, default=400, help="c3 GMRES iteration cap")
    ap.add_argument("--restart", type=int, default=30, help="c3 GMRES restart (the configuration: 30)")
    ap.add_argument("--checkpoint-every", type=int, default=None,
                    help="c5 trajectory layout: -2/-1 tape with/without face states, 1 every "
                         "state, k>1 checkpoints, -2-c binomial schedule on c state slots, 0 "
                         "the richest that fits; default the binomial schedule on ceil(sqrt(steps)) "
                         "slots")
    ap.add_argument("--exact-limiter", action="store_true",
                    help="c5: differentiate the limiter through its active branch (default: frozen)")
    ap.add_argument("--no-breakdown", action="store_true",
                    help="c5: skip the comparison layouts (full tape, no stretch tape, other limiter mode)")
    ap.add_argument("--horizons", default="1,2,5,10,25,50,100,200,500", help="c5fd: trajectory lengths")
    ap.add_argument("--fd-steps", default="1e-6", help="c5fd: central-difference steps")
    ap.add_argument("--directions", type=int, default=64, help="c5fd: random directions")
    ap.add_argument("--frozen-map", action="store_true",
                    help="c5fd: also difference the frozen-limiter map for the frozen-limiter gradient")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: c2 200, c3 10, c4 300, c5 500, c3adj 10)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=1000)
    ap.add_argument("--order", default="rcm", choices=["hilbert", "rcm", "none"],
                    help="cell renumbering of the mesh service")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="C2: keep the default launch shape (256 x 1) instead of the autotuned one")
    ap.add_argument("--launch", default=None, metavar="BSxCPT",
                    help="C2: this launch shape (workgroup size x cells per thread), no autotuning")
    ap.add_argument("--time-integration", default="splitting", choices=["splitting", "imex", "implicit"],
                    help="c2 strategy: splitting (the headline), imex (ARK2), implicit (BDF2 Newton-Krylov)")
    ap.add_argument("--no-limiter", dest="limiter", action="store_false",
                    help="c2 with imex / implicit: unlimited reconstruction")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d refused)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.steps is None:
        args.steps = {"c1": 200, "c2": 200, "c3": 10, "c4": 300, "c5": 500, "c5fd": 0, "c3adj": 10}[args.workload]
    zdm = load_package()
    import zoomy_dmplex_amd.cases as cases
    if args.workload == "c1":
        return run_c1(zdm, cases, args)
    if args.workload == "c5":
        return run_c5(zdm, cases, args)
    if args.workload == "c5fd":
        return run_c5_fd(zdm, cases, args)
    if args.workload == "c3":
        return run_c3(zdm, cases, args)
    if args.workload == "c3adj":
        return run_c3adj(zdm, cases, args)
    if args.workload == "c4":
        return run_c4(zdm, cases, args)
    if args.time_integration != "splitting":
        return run_c2_time_integration(zdm, cases, args)

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
    # launch shape of the transport kernels by measurement (the autotuner restores the state)
    shape, tune = (256, 1), {}
    if args.launch:
        shape = tuple(int(v) for v in args.launch.lower().split("x"))
        s.set_launch(*shape)
    elif args.autotune:
        bs, cpt, tune = s.autotune(3)
        shape = (bs, cpt)
    s.synchronize()

    # kernel events on every TIMING_STRIDE-th step of the timed region: live per-kernel times
    # without a marker between the kernels of every step
    s.enable_timing(TIMING_STRIDE)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    flux_ms, flux_n = s.kernel_time(1)
    prep_ms, prep_n = s.kernel_time(0)
    post_ms, post_n = s.kernel_time(2)
    fpost_ms, fpost_n = s.kernel_time(3)
    s.enable_timing(False)
    t_sim, dt, nsteps = s.time()
    Xg, _ = s.download()
    if not np.all(np.isfinite(Xg)):
        raise RuntimeError("non-finite state after the benchmark")
    nontemporal = s.flux_nontemporal() and shape[1] == 1   # the by-size default the run used

    ncells = mesh.ncells
    F = mesh.nfaces / ncells
    value = ncells * args.steps / elapsed / 1e6
    if fpost_n:
        # the step's two flux passes: k_flux (stage) and k_flux_post (final stage + post-step)
        nl = flux_n + fpost_n
        bpc = (flux_n * flux_pass_bytes_per_cell(info.n_q, F) +
               fpost_n * final_post_bytes_per_cell(info.n_q, F)) / nl
        launch_ms = (flux_n * flux_ms + fpost_n * fpost_ms) / nl
        rkernel = "k_flux (stage: flux + fluctuations gather + SSP stage update) and k_flux_post (final " \
                  "stage fused with the post-step: constraints, Newton source, constraints, CFL)"
        traffic = pmc_traffic("k_flux+k_flux_post", ncells)
        tsrc = "rocprofv3 FETCH_SIZE + WRITE_SIZE, mean of the k_flux and k_flux_post launches, " \
               "copy-kernel calibrated: bytes past the L2, served by the Infinity Cache or HBM, so an " \
               "upper bound on HBM bytes (profiles/k_flux_c2_2M_pmc.md)"
    else:
        nl, bpc, launch_ms = flux_n, flux_launch_bytes_per_cell(info.n_q, F), flux_ms
        rkernel = "k_flux (flux + fluctuations gather + SSP stage update)"
        traffic = pmc_traffic("k_flux", ncells)
        tsrc = "rocprofv3 FETCH_SIZE + WRITE_SIZE, mean of the stage and final launches, copy-kernel " \
               "calibrated (profiles/k_flux_c2_2M_pmc.md)"
    achieved = bpc * ncells / (launch_ms * 1e-3) / 1e9
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
            "flux_nontemporal": nontemporal,
            "n_q": info.n_q,
            "faces_per_cell": round(F, 4),
            "step_bytes_per_cell": round(step_bytes_per_cell(info.n_q, F), 1),
            "kernel_ms": {"prep": round(prep_ms, 4), "flux": round(flux_ms, 4),
                          "post": round(post_ms, 4), "flux_post": round(fpost_ms, 4)},
            "kernel_ms_sample": "HIP events on every %d-th timed step (%d flux, %d flux_post launches)" % (
                TIMING_STRIDE, flux_n, fpost_n),
            "launch": {"block_size": shape[0], "cells_per_thread": shape[1],
                       "autotune_ms_per_step": {"%dx%d" % k: round(v, 4) for k, v in tune.items()}},
            "sim_time": t_sim,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": rkernel,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": tsrc,
            "bytes_per_cell": round(bpc, 2),
            "launches": nl,
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
