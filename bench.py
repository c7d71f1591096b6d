#!/usr/bin/env python3
"""Benchmark of the Godunov step (HLLC, perfect gas) at 512^3 on one MI355X.

A "step" is one godunov update (godunov.hpp:15-92) over the 512^3 interior of
514^3 arrays, FP64, int64 indexing, on a seeded random state (rho ~ U[0.5,2],
u_k ~ N(0,0.3), p ~ U[0.5,2], gamma = 1.4, dx = 1/512; dt = 0.5 * time_step).
`value` is interior cell-updates per second of that step with all inputs resident
in HBM. The JSON line also carries the roofline of the dominant kernel (the tiled
godunov kernel, algorithmic bytes s*(5*N_arr + 10*N_int) per launch), the CPU
oracle timed on the host beside it, and secondary figures (fused full step,
conversions, time step, FP32, reference-order kernel) under "extra".

    python bench.py [--steps K] [--warmup W] [--n 512] [--no-cpu] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s & % HBM peak, 512^3 Godunov step; batched & adjoint vs plain"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512, help="interior cells per axis")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline leg")
    ap.add_argument("--quick", action="store_true", help="headline godunov only (for profiling)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-fp32-report", action="store_true", help="skip the FP32-vs-FP64 implosion report")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 4 GiB device copy (HBM counter calibration)")
    ap.add_argument("--config", type=str, default=None, help="godunov 'variant,by,kchunk'")
    ap.add_argument("--batch-only", type=int, default=0, metavar="STEPS",
                    help="profiling aid: only the batched ensemble, STEPS graph-replayed steps after autotuning")
    ap.add_argument("--adjoint-only", type=int, default=0, metavar="RUNS",
                    help="profiling aid: only the adjoint workload's forward + backward, RUNS timed runs after "
                         "two warm-up runs")
    ap.add_argument("--f32-only", type=int, default=0, metavar="STEPS",
                    help="profiling aid: only the FP32 godunov step at n^3 (STEPS launches after warm-up), then "
                         "three FP32 prim_to_cons launches on the same arrays (counter calibration)")
    ap.add_argument("--traffic-record", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"), default=None,
                    help="write the profiled-traffic record from two rocprofv3 --pmc counter CSVs and exit")
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="with --quick: also run prim_to_cons on the same arrays (HBM counter calibration: "
                         "one 8-byte load and one 8-byte store per element, no reuse)")
    return ap.parse_args()


class Timer:
    """HIP events on the stream the kernels are launched on."""

    def __init__(self, torch, stream):
        self.torch = torch
        self.stream = stream

    def run(self, fn, reps: int) -> float:
        torch = self.torch
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(self.stream)
        for _ in range(reps):
            fn()
        end.record(self.stream)
        end.synchronize()
        return start.elapsed_time(end) / reps  # ms per launch


def cpu_baseline(n: int):
    """The reference's CPU path stood in for by the oracle (a reference-order C++
    restatement, OpenMP; Kokkos is unavailable offline), built for timing as the
    reference's CPU build would be (g++ -O3 -march=native -fopenmp, default
    contraction; compiled here, on the host that runs it) and run in a subprocess
    with OMP_PROC_BIND=close, OMP_PLACES=cores (setups/adastra/genoa/job.sh:20-24)
    on the threads this process may use. Every operator at n^3 in FP64 and FP32 and
    the 64^3 x 100-step implosion, medians of 3 (oracle/timing.py)."""
    import subprocess

    odir = ROOT / "oracle"
    lib, build = odir / "liboracle_native.so", "native (-O3 -march=native -fopenmp)"
    try:
        # -B: always for this host's ISA, never a library built elsewhere
        subprocess.run(["make", "-s", "-B", "-C", str(odir), "native"], check=True, capture_output=True, timeout=300)
    except (subprocess.SubprocessError, OSError) as e:
        log(f"native oracle build failed ({e}); timing the parity build")
        lib, build = odir / "liboracle.so", "parity (-O3 -march=x86-64-v3 -ffp-contract=off)"
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))
    env["OMP_PROC_BIND"] = "close"
    env["OMP_PLACES"] = "cores"
    r = subprocess.run([sys.executable, str(odir / "timing.py"), "--n", str(n), "--lib", str(lib), "--build", build],
                       env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        log(f"CPU timing failed: {r.stderr[-2000:]}")
        return None
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    g = rec["f64"]["godunov"]
    rnd = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}  # noqa: E731
    return {
        "value": round(g["Mcell_updates_per_s"], 2),
        "unit": "Mcell-updates/s",
        "cores": rec["threads"],
        "kind": "port",
        "sample": f"godunov FP64 at {n}^3 interior ({n + 2}^3 arrays), seeded random state (seed 1234), median "
                  f"of 3; oracle = reference-order C++ restatement, OpenMP (Kokkos unavailable offline), build "
                  f"{rec['build']}, {rec['threads']} threads, OMP_PROC_BIND=close OMP_PLACES=cores",
        "cpu_model": rec["cpu_model"],
        "nproc": rec["nproc"],
        "cpus_available": rec["cpus_available"],
        "operators_f64": {k: rnd(v) for k, v in rec["f64"].items()},
        "operators_f32": {k: rnd(v) for k, v in rec["f32"].items()},
        "implosion_64_fp64_100_steps": rnd(rec["implosion_64"]),
        "seconds": round(rec["seconds"], 1),
    }


def adjoint_leg(torch, eb, ex, timer, n: int = 128, nsteps: int = 200, k: int = 20, ndirs: int = 64):
    """Forward + hand-written adjoint of a 200-step FP64 rollout of the diagonal
    implosion at 128^3 (density perturbed by N(0,0.01) * bump, seed 1234), loss =
    sum rho^2 over the octant x, y, z < 0.25, checkpoints every 20 steps, against
    the plain forward step loop (fused step, hipGraph) over the same steps; then
    central FD (eps = 1e-6, dt frozen) on `ndirs` density directions (seeds 5678+j)."""
    import statistics

    G = 1.4
    eos = eb.PerfectGas(G)
    mesh = eb.UniformMesh3d(1.0 / n, 1.0 / n, 1.0 / n)
    P0 = eb.workloads.adjoint_state(ex, n)
    c = (torch.arange(n + 2, dtype=torch.float64, device=P0.d.device) - 0.5) / n
    inner = (c > 0) & (c < 0.25)
    mask = inner[:, None, None] & inner[None, :, None] & inner[None, None, :]

    def clone_prim():
        return eb.EulerPrimArrays(*(t.clone() for t in P0.tensors()), index_bits=P0.index_bits)

    # plain forward: the fused step loop, graph-replayed and eager; the faster of
    # the two is the denominator of the ratio
    plain = {}
    for use_graph in (True, False):
        prim = clone_prim()
        cons = eb.EulerConsArrays(*(torch.empty_like(t) for t in P0.tensors()), index_bits=P0.index_bits)
        eb.prim_to_cons(ex, prim, cons, eos)
        sim = eb.simulation.Simulation(ex, prim, cons, eos, mesh, 0.5)
        sim.run(4, use_graph=use_graph)
        sim.run(nsteps, use_graph=use_graph)
        plain["graph" if use_graph else "eager"] = statistics.median(
            timer.run(lambda: sim.run(nsteps, use_graph=use_graph), 1) for _ in range(3))
        del sim, prim, cons
    ms_plain = min(plain.values())

    def wall(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3, r

    def fwd_only():
        # the differentiable forward (checkpoints and the last segment kept), no backward
        ts = [t.clone().requires_grad_(True) for t in P0.tensors()]
        return eb.adjoint.rollout(eb.EulerPrimArrays(*ts, index_bits=P0.index_bits), eos, mesh, 0.5, nsteps, k,
                                  exec_space=ex)

    def fwd_bwd():
        ts = [t.clone().requires_grad_(True) for t in P0.tensors()]
        tape = []
        out = eb.adjoint.rollout(eb.EulerPrimArrays(*ts, index_bits=P0.index_bits), eos, mesh, 0.5, nsteps, k,
                                 exec_space=ex, tape=tape)
        (out[0] * out[0] * mask).sum().backward()
        return [t.grad for t in ts], tape[0].dts.clone()

    # two warm-up rounds (first-touch allocations of the caching allocator), then
    # the median of three timed runs of each
    for _ in range(2):
        fwd_only()
        fwd_bwd()
    ms_fwd = statistics.median(wall(fwd_only)[0] for _ in range(3))
    runs = [wall(fwd_bwd) for _ in range(3)]
    ms_total = statistics.median(r[0] for r in runs)
    grads, dts = runs[-1][1]

    def loss(rho):
        with torch.no_grad():
            P = eb.EulerPrimArrays(rho, *(t.clone() for t in P0.tensors()[1:]), index_bits=P0.index_bits)
            o = eb.adjoint.rollout(P, eos, mesh, 0.5, nsteps, nsteps, exec_space=ex, frozen_dts=dts)
            return float((o[0] * o[0] * mask).sum().item())

    # Central differences with dt frozen; the one-sided differences tell a kink
    # inside [-eps, eps] (fd+ != fd-, the adjoint then equals one of them) from the
    # rounding of the 200-step loss (~1e-4 absolute in fd at eps 1e-6).
    eps = 1e-6
    errs, abs_errs, within, between = [], [], 0, 0
    t0 = time.perf_counter()
    l0 = loss(P0.d.clone())
    for j in range(ndirs):
        v = eb.workloads.density_direction(n, j, P0.d.device)
        lp, lm = loss(P0.d + eps * v), loss(P0.d - eps * v)
        fd, fp, fm = (lp - lm) / (2 * eps), (lp - l0) / eps, (l0 - lm) / eps
        an = float((grads[0] * v).sum())
        allow = 3e-6 * abs(an) + 2e-4
        errs.append(abs(fd - an) / max(abs(an), 1e-300))
        abs_errs.append(abs(fd - an))
        within += abs(fd - an) <= allow
        between += min(fp, fm) - allow <= an <= max(fp, fm) + allow
    fd_s = time.perf_counter() - t0
    return {
        "workload": f"{nsteps}-step FP64 rollout, diagonal implosion {n}^3, rho += N(0,0.01)*bump (seed 1234), "
                    f"loss sum rho^2 over octant x,y,z<0.25, checkpoints every {k}",
        "plain_forward_ms": round(ms_plain, 2),
        "plain_forward_graph_ms": round(plain["graph"], 2),
        "plain_forward_eager_ms": round(plain["eager"], 2),
        "rollout_forward_ms": round(ms_fwd, 2),
        "forward_plus_adjoint_ms": round(ms_total, 2),
        "timing": "median of 3 runs after 2 warm-up rounds; plain forward: the faster of the graph-replayed "
                  "and the eager fused-step loop",
        "ratio_vs_plain_forward": round(ms_total / ms_plain, 3),
        "ratio_vs_rollout_forward": round(ms_total / ms_fwd, 3),
        "fd_check": {"directions": ndirs, "eps": eps, "field": "rho (N(0,1)*bump, seeds 5678+j)",
                     "loss": l0, "max_rel_err": float(max(errs)), "median_rel_err": float(statistics.median(errs)),
                     "max_abs_err": float(max(abs_errs)), "n_below_1e-6_rel": sum(e < 1e-6 for e in errs),
                     "bar": "|fd - <grad,v>| <= 3e-6 |<grad,v>| + 2e-4 (rounding floor of the loss / eps)",
                     "n_within_bar": int(within), "n_between_one_sided": int(between),
                     "seconds": round(fd_s, 1)},
    }


def stream_probe(torch, eb, ex, timer, n_arr: int, dtype):
    """eb_stream_probe_*: b.f += a.f for the five fields over n_arr cells (five
    fields read from one array set, five read and written in place in another, one
    access of the element size each), the plain flux kernel's HBM traffic per cell
    with no tiles, halos or barriers. Returns (ms per launch, GB/s)."""
    import ctypes

    from benchmark_euler_kernels_amd import operators as ebo
    from benchmark_euler_kernels_amd._lib import call

    suf = "f64" if dtype == torch.float64 else "f32"
    a = eb.create_prim_arrays_1d(ex, n_arr, dtype)
    b = eb.create_cons_arrays_1d(ex, n_arr, dtype)
    for t in (*a.tensors(), *b.tensors()):
        t.zero_()
    pa, pb = ebo.prim_ptrs(a), ebo.cons_ptrs(b)
    fn = lambda: call(f"eb_stream_probe_{suf}", ex.handle, ctypes.byref(pa), ctypes.byref(pb), n_arr)  # noqa: E731
    timer.run(fn, 3)
    ms = timer.run(fn, 10)
    elem = 8 if dtype == torch.float64 else 4
    return ms, 15 * elem * n_arr / ms / 1e6


def batch_leg(torch, eb, ex, timer, count: int = 256, nsteps: int = 500, base_steps: int = 50):
    """The measured ensemble: `count` diagonal implosions (instance i drawn from
    seed 1234+i: interior extents U{32..160} per axis, gamma ~ U[1.2,1.67],
    interface offset ~ U[0.3,0.6], inner rho, p ~ U[0.05,0.3]), FP32, CFL 0.5,
    per-instance dt on the device. One persistent launch per step (autotuned
    tiles, hipGraph-replayed) against one fused launch per instance and step."""
    prims, conss, eoss, meshes, cfls, params = eb.batch.random_ensemble(ex, count, seed=1234, dtype=torch.float32,
                                                                        index_bits=32, cfl=0.5)
    cells = sum(p["n"][0] * p["n"][1] * p["n"][2] for p in params)
    arr_cells = sum((p["n"][0] + 2) * (p["n"][1] + 2) * (p["n"][2] + 2) for p in params)
    # fused step per instance: primitives read on every array cell, conservatives read
    # and written and next primitives written on the interior cells (FP32)
    step_bytes = 4 * (5 * arr_cells + 15 * cells)
    ens = eb.batch.Ensemble(ex, prims, conss, eoss, meshes, cfls)
    ens.start()
    ens.step()
    t0 = time.perf_counter()
    tune_ms = ens.autotune(reps=3)
    tune_s = time.perf_counter() - t0
    tiles = ens.tiles()
    ens.run(3, use_graph=True)  # capture
    ms_graph = timer.run(lambda: ens.run(nsteps - 4, use_graph=True), 1) / (nsteps - 4)
    ms_eager = timer.run(lambda: ens.run(base_steps), 1) / base_steps
    ens.run(2, use_graph=True, per_instance=True)  # capture the per-instance pair
    ms_pi_graph = timer.run(lambda: ens.run(base_steps, use_graph=True, per_instance=True), 1) / base_steps
    ms_pi_eager = timer.run(lambda: ens.run(base_steps, per_instance=True), 1) / base_steps
    # the single-grid kernel with linear tiles as well, so the baseline is the best per-instance form
    cfg = eb.get_godunov_config()
    eb.set_godunov_config(3, 8, 0)
    ens.per_instance_graph = None
    ens.run(2, use_graph=True, per_instance=True)
    ms_pi_lin = timer.run(lambda: ens.run(base_steps, use_graph=True, per_instance=True), 1) / base_steps
    eb.set_godunov_config(*cfg)
    _, t = ens.state()
    steps_done = ens.steps_done
    ens.close()
    rate = lambda ms: round(cells / (ms * 1e-3) / 1e6, 1)  # noqa: E731
    return {
        "workload": f"{count} diagonal implosions, interior extents U{{32..160}} per axis (seed 1234+i), "
                    f"gamma U[1.2,1.67], offset U[0.3,0.6], inner rho,p U[0.05,0.3], FP32, CFL 0.5",
        "interior_cells": cells,
        "autotuned_tiles": tiles,
        "autotune_ms_per_launch": round(tune_ms, 4),
        "autotune_seconds": round(tune_s, 2),
        "batched_graph_ms_per_step": round(ms_graph, 4),
        "batched_graph_Mcell_updates_per_s": rate(ms_graph),
        "algorithmic_bytes_per_step": step_bytes,
        "batched_graph_GBps_algorithmic": round(step_bytes / ms_graph / 1e6, 1),
        "batched_graph_hbm_frac": round(step_bytes / ms_graph / 1e6 / 8000.0, 3),
        "batched_eager_ms_per_step": round(ms_eager, 4),
        "per_instance_eager_ms_per_step": round(ms_pi_eager, 4),
        "per_instance_eager_Mcell_updates_per_s": rate(ms_pi_eager),
        "per_instance_graph_ms_per_step": round(ms_pi_graph, 4),
        "per_instance_graph_Mcell_updates_per_s": rate(ms_pi_graph),
        "per_instance_linear_tiles_graph_ms_per_step": round(ms_pi_lin, 4),
        "per_instance_linear_tiles_graph_Mcell_updates_per_s": rate(ms_pi_lin),
        "speedup_vs_per_instance_eager": round(ms_pi_eager / ms_graph, 3),
        "speedup_vs_per_instance_graph": round(ms_pi_graph / ms_graph, 3),
        "speedup_vs_best_per_instance": round(min(ms_pi_eager, ms_pi_graph, ms_pi_lin) / ms_graph, 3),
        "steps_done": steps_done,
        "simulated_time_min_max": [float(t.min()), float(t.max())],
        "timed_steps": {"batched_graph": nsteps - 4, "others": base_steps},
    }


def adjoint_profile(torch, eb, ex, runs: int, n: int = 128, nsteps: int = 200, k: int = 20):
    """The adjoint leg's differentiable forward and backward alone (the workload of
    adjoint_leg, without the plain forward and the FD check), for counters."""
    import statistics

    eos = eb.PerfectGas(1.4)
    mesh = eb.UniformMesh3d(1.0 / n, 1.0 / n, 1.0 / n)
    P0 = eb.workloads.adjoint_state(ex, n)
    c = (torch.arange(n + 2, dtype=torch.float64, device=P0.d.device) - 0.5) / n
    inner = (c > 0) & (c < 0.25)
    mask = inner[:, None, None] & inner[None, :, None] & inner[None, None, :]

    def fwd_bwd():
        ts = [t.clone().requires_grad_(True) for t in P0.tensors()]
        out = eb.adjoint.rollout(eb.EulerPrimArrays(*ts, index_bits=P0.index_bits), eos, mesh, 0.5, nsteps, k,
                                 exec_space=ex)
        (out[0] * out[0] * mask).sum().backward()

    for _ in range(2):
        fwd_bwd()
    ms = []
    for _ in range(runs):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fwd_bwd()
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    return {"forward_plus_adjoint_ms_median": round(statistics.median(ms), 2), "runs": runs}


def batch_profile(torch, eb, ex, timer, nsteps: int, count: int = 256):
    """The measured ensemble stepped nsteps times after autotuning (for counters)."""
    prims, conss, eoss, meshes, cfls, params = eb.batch.random_ensemble(ex, count, seed=1234, dtype=torch.float32,
                                                                        index_bits=32, cfl=0.5)
    ens = eb.batch.Ensemble(ex, prims, conss, eoss, meshes, cfls)
    ens.start()
    ens.step()
    ens.autotune(reps=3)
    ens.run(3, use_graph=True)
    ms = timer.run(lambda: ens.run(nsteps, use_graph=True), 1) / nsteps
    tiles = ens.tiles()
    ens.close()
    return {"batched_graph_ms_per_step": round(ms, 4), "tiles": tiles, "steps": nsteps}


def fp32_report(torch, eb, ex, n: int = 512, nsteps: int = 1000, cfl: float = 0.5):
    """Diagonal implosion (x+y+z > 0.45: rho = p = 1, else rho = 0.125, p = 0.14;
    u = 0; gamma 1.4; unit cube) at n^3, nsteps steps, FP32 state and arithmetic
    against FP64 (the FP64 GPU path, bitwise equal to the oracle). Two FP32 runs:
    (a) its own CFL time steps (fused step loop, graph), (b) driven by the FP64
    run's time steps rounded to FP32 (reference-order loop), so that both end at
    the same simulated time. Errors of rho, p and u_k over the interior at t_end;
    conservation residuals from FP64 compensated totals and boundary outflows."""
    eos = eb.PerfectGas(1.4)
    mesh = eb.UniformMesh3d(1.0 / n, 1.0 / n, 1.0 / n)

    def state(dtype):
        prim = eb.workloads.diagonal_implosion(ex, n, dtype, index_bits=64)
        cons = eb.EulerConsArrays(*(torch.empty_like(t) for t in prim.tensors()), index_bits=64)
        eb.prim_to_cons(ex, prim, cons, eos)
        return prim, cons

    def relres(cons_report):
        out = {}
        for name, r in cons_report.items():
            ref = max(abs(r["total0"]), abs(r["total"]), abs(r["outflow"]))
            out[name] = {"residual": r["residual"], "relative": r["residual"] / ref if ref else 0.0}
        return out

    # FP64 reference run, eager, recording the raw dt of every step.
    t0 = time.perf_counter()
    prim, cons = state(torch.float64)
    sim = eb.simulation.Simulation(ex, prim, cons, eos, mesh, cfl, track_conservation=True)
    tape = torch.empty(nsteps, dtype=torch.float64, device=prim.d.device)
    sim.start()
    for k in range(nsteps):
        tape[k:k + 1].copy_(sim.last_dt())
        sim.step()
    sim.finish()
    t64 = float(sim.time.item())
    cons64 = relres(sim.conservation())
    ref = [t.clone() for t in prim.tensors()]
    del sim, prim, cons
    torch.cuda.empty_cache()

    def errors(p32):
        inner = (slice(1, -1),) * 3
        out = {}
        for name, a, b in zip(("rho", "p", "u0", "u1", "u2"), p32.tensors(), ref):
            d = (a[inner].double() - b[inner]).abs()
            scale = b[inner].abs().mean().item()
            out[name] = {"L1": d.mean().item(), "Linf": d.max().item(),
                         "L1_relative": d.mean().item() / scale if scale else None}
        return out

    # (a) FP32 with its own time steps.
    prim, cons = state(torch.float32)
    sim = eb.simulation.Simulation(ex, prim, cons, eos, mesh, cfl, track_conservation=True)
    sim.run(nsteps, use_graph=True)
    sim.finish()
    t32 = float(sim.time.item())
    own = {"t_end": t32, "errors": errors(prim), "conservation": relres(sim.conservation())}
    del sim, prim, cons
    torch.cuda.empty_cache()

    # (b) FP32 driven by the FP64 time steps (reference-order loop, dt on device).
    prim, cons = state(torch.float32)
    tape32 = tape.to(torch.float32)
    hllc = eb.hllc()
    t_b = 0.0
    for k in range(nsteps):
        eb.godunov(ex, prim, cons, eos, mesh, hllc, tape32[k:k + 1], cfl=cfl)
        eb.cons_to_prim(ex, cons, prim, eos)
    t_b = float((tape32.double() * float(np.float32(cfl))).sum().item())
    driven = {"t_end": t_b, "errors": errors(prim)}
    del prim, cons, ref
    torch.cuda.empty_cache()
    return {
        "workload": f"diagonal implosion {n}^3 (x+y+z>0.45), gamma 1.4, CFL {cfl}, {nsteps} steps",
        "fp64": {"t_end": t64, "conservation": cons64},
        "fp32_own_dt": own,
        "fp32_fp64_dt": driven,
        "seconds": round(time.perf_counter() - t0, 1),
    }


TRAFFIC_RECORD = ROOT / "profiles" / "godunov_f64_512_traffic.json"


def source_digest() -> str:
    """SHA-256 of the library's kernel and C ABI sources: a traffic figure measured
    on other sources is not reported as this build's."""
    import hashlib

    h = hashlib.sha256()
    csrc = ROOT / "benchmark-euler-kernels_amd" / "csrc"
    for p in sorted(list(csrc.glob("*.hpp")) + list(csrc.glob("*.hip")) + [ROOT / "include" / "euler_capi.h"]):
        h.update(p.name.encode())
        h.update(p.read_bytes())
    return h.hexdigest()


def profiled_traffic(kernel: str, n: int):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    calibrated on prim_to_cons) for this kernel and workload, from the committed
    record; None when the record is for another kernel, size or source tree."""
    if n != 512 or not TRAFFIC_RECORD.exists():
        return None
    rec = json.loads(TRAFFIC_RECORD.read_text())
    if rec.get("kernel") != kernel or rec.get("library_source_sha256") != source_digest():
        log("profiled traffic record is for another kernel or source tree: traffic not reported")
        return None
    return rec["traffic_bytes_per_launch"]


def write_traffic_record(fetch_csv: str, write_csv: str, n: int = 512) -> dict:
    """Traffic record from two rocprofv3 --pmc runs of `bench.py --quick
    --traffic-calibration` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch). Each
    counter is scaled by what it reads for k_prim_to_cons<double, long> on the same
    arrays, which loads and stores each of 5 fields once, 8 bytes per element."""
    import csv
    import statistics

    def per_kernel(path, counter):
        out = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]) * 1024.0)
        return out

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    na = n + 2
    calib_bytes = 5 * 8 * na**3

    def pick(d, pat):
        keys = [k for k in d if pat in k]
        assert len(keys) == 1, (pat, keys)
        return keys[0], statistics.median(d[keys[0]])

    gname, gf = pick(fetch, "k_godunov_pair<double, long")
    _, gw = pick(write, "k_godunov_pair<double, long")
    _, cf = pick(fetch, "k_prim_to_cons<double, long>")
    _, cw = pick(write, "k_prim_to_cons<double, long>")
    fs, ws = calib_bytes / cf, calib_bytes / cw
    kernel = "eb::" + gname.split("eb::", 1)[1].split("(", 1)[0]
    rec = {
        "kernel": kernel,
        "workload": f"godunov FP64 {n}^3 interior, int64 indexing",
        "traffic_bytes_per_launch": gf * fs + gw * ws,
        "fetch_bytes": gf * fs,
        "write_bytes": gw * ws,
        "algorithmic_bytes": 8 * (5 * na**3 + 10 * n**3),
        "fetch_scale": fs,
        "write_scale": ws,
        "calibration": "k_prim_to_cons<double, long> on the same (n+2)^3 arrays: 5 fields read and written once",
        "library_source_sha256": source_digest(),
    }
    TRAFFIC_RECORD.write_text(json.dumps(rec, indent=1) + "\n")
    return rec


def kernel_name(eb) -> str:
    """The FP64 flux kernel the current configuration launches (int64 indices)."""
    variant = eb.get_godunov_config()[0]
    by = eb.godunov_tile_height(fused=False)
    kind = {0: "k_godunov_faces", 1: "k_godunov_tiled", 2: "k_godunov_stream", 3: "k_godunov_linear",
            4: "k_godunov_pair"}[variant]
    return f"eb::{kind}<double, long, {by}, false>"


def main():
    args = parse()
    if args.traffic_record:
        print(json.dumps(write_traffic_record(*args.traffic_record)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(the multi-GPU halo design is documented, not built)"}))
        sys.exit(2)

    import torch

    from __graft_entry__ import load_package

    eb = load_package()
    if args.config:
        eb.set_godunov_config(*(int(v) for v in args.config.split(",")))
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    ex = eb.ExecSpace(dev)
    timer = Timer(torch, ex.stream)
    if args.adjoint_only:
        log(f"adjoint workload only: {args.adjoint_only} forward + backward runs")
        print(json.dumps(adjoint_profile(torch, eb, ex, args.adjoint_only)), flush=True)
        ex.close()
        return
    if args.batch_only:
        log(f"batched ensemble only: {args.batch_only} graph-replayed steps")
        print(json.dumps(batch_profile(torch, eb, ex, timer, args.batch_only)), flush=True)
        ex.close()
        return
    n = args.n
    na = n + 2
    n_int = n**3
    n_arr = na**3
    gamma = 1.4
    eos = eb.PerfectGas(gamma)
    mesh = eb.UniformMesh3d(1.0 / n, 1.0 / n, 1.0 / n)

    def alloc(dtype):
        prim = eb.to_mdspan(eb.create_prim_arrays_1d(ex, n_arr, dtype), na, na, na, index_bits=64)
        cons = eb.to_mdspan(eb.create_cons_arrays_1d(ex, n_arr, dtype), na, na, na, index_bits=64)
        eb.init_random(ex, prim, seed=1234)
        eb.prim_to_cons(ex, prim, cons, eos)
        return prim, cons

    if args.f32_only:
        log(f"FP32 godunov only: {args.f32_only} launches at {n}^3")
        prim32, cons32 = alloc(torch.float32)
        dt32 = 0.5 * eb.time_step(ex, prim32, eos, mesh)
        hllc32 = eb.hllc()
        timer.run(lambda: eb.godunov(ex, prim32, cons32, eos, mesh, hllc32, dt32), 3)
        ms32 = timer.run(lambda: eb.godunov(ex, prim32, cons32, eos, mesh, hllc32, dt32), args.f32_only)
        for _ in range(3):
            eb.prim_to_cons(ex, prim32, cons32, eos)
        torch.cuda.synchronize()
        bytes32 = 4 * (5 * n_arr + 10 * n_int)
        print(json.dumps({"godunov_f32_ms": round(ms32, 4), "algorithmic_bytes": bytes32,
                          "GBps_algorithmic": round(bytes32 / ms32 / 1e6, 1), "launches": args.f32_only}), flush=True)
        ex.close()
        return

    log(f"allocating {n}^3 FP64 state")
    prim, cons = alloc(torch.float64)
    dt = 0.5 * eb.time_step(ex, prim, eos, mesh)
    hllc = eb.hllc()

    def step():
        eb.godunov(ex, prim, cons, eos, mesh, hllc, dt)

    log(f"warmup {args.warmup}")
    timer.run(step, max(args.warmup, 1))
    torch.cuda.synchronize()
    log(f"timing {args.steps} godunov steps")
    t0 = time.perf_counter()
    ms = timer.run(step, args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_wall = wall * 1e3 / args.steps
    s = 8
    bytes_godunov = s * (5 * n_arr + 10 * n_int)
    achieved = bytes_godunov / (ms * 1e-3) / 1e9
    value = n_int / (ms_wall * 1e-3) / 1e6
    log(f"godunov FP64: {ms:.3f} ms/launch (events), {ms_wall:.3f} ms/step (wall), {achieved:.0f} GB/s")

    if args.traffic_calibration:
        for _ in range(3):
            eb.prim_to_cons(ex, prim, cons, eos)
        torch.cuda.synchronize()

    extra = {
        "godunov_f64_ms_event": round(ms, 4),
        "godunov_f64_Mcells_per_s_ref_convention": round(n_arr / (ms * 1e-3) / 1e6, 1),
        "godunov_config": list(eb.get_godunov_config()),
    }

    if not args.quick:
        # the same launches timed one by one (an event pair around each), for the
        # median and the minimum
        import statistics

        per = [timer.run(step, 1) for _ in range(args.steps)]
        extra["godunov_f64_per_launch_ms"] = {"median": round(statistics.median(per), 4),
                                              "min": round(min(per), 4), "max": round(max(per), 4),
                                              "launches": len(per)}

        reps = max(3, args.steps // 2)
        # Streaming probe with the flux kernel's read:write mix (two arrays read per
        # array written): c = a + b over three 1 GiB FP64 arrays, i.e. what a plain
        # streaming kernel with that mix reaches on this part.
        pa = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        pb = torch.ones_like(pa)
        pc = torch.empty_like(pa)
        with torch.cuda.stream(ex.stream):
            timer.run(lambda: torch.add(pa, pb, out=pc), 3)
            ms_probe = timer.run(lambda: torch.add(pa, pb, out=pc), 10)
        probe = 3 * pa.numel() * 8 / ms_probe / 1e6
        extra["stream_probe_read2_write1"] = {"kernel": "torch.add(a, b, out=c) over three 1 GiB FP64 arrays",
                                              "GBps": round(probe, 1),
                                              "godunov_f64_frac_of_probe": round(achieved / probe, 4)}
        del pa, pb, pc
        # The flux kernel's own traffic per cell streamed with no arithmetic, tiles,
        # halos or barriers: what plain streaming of that traffic reaches on this part.
        ms_sprobe, sprobe_gbs = stream_probe(torch, eb, ex, timer, n_arr, torch.float64)
        extra["stream_probe_godunov_traffic"] = {
            "kernel": "eb_stream_probe_f64: b.f += a.f for 5 fields over 514^3 FP64 arrays (8-byte accesses)",
            "ms": round(ms_sprobe, 4), "GBps": round(sprobe_gbs, 1), "hbm_frac": round(sprobe_gbs / HBM_PEAK_GBS, 4),
            "godunov_f64_frac_of_probe": round(achieved / sprobe_gbs, 4)}
        # Reference-order kernel (six solves per cell), same state.
        ms_ro = timer.run(lambda: eb.godunov(ex, prim, cons, eos, mesh, hllc, dt, reference_order=True), reps)
        extra["godunov_reference_order_f64_ms"] = round(ms_ro, 4)
        # Pointwise operators and the reduction (n^3 conventions of the reference
        # benchmarks use (n+2)^3 arrays here; bytes counted on the arrays timed).
        ms_c2p = timer.run(lambda: eb.cons_to_prim(ex, cons, prim, eos), reps)
        ms_p2c = timer.run(lambda: eb.prim_to_cons(ex, prim, cons, eos), reps)
        dt_dev = torch.empty(1, dtype=torch.float64, device=dev)
        ms_ts = timer.run(lambda: eb.time_step_async(ex, prim, eos, mesh, out=dt_dev), reps)
        extra["cons_to_prim_f64"] = {"ms": round(ms_c2p, 4), "GBps": round(10 * s * n_arr / ms_c2p / 1e6, 1)}
        extra["prim_to_cons_f64"] = {"ms": round(ms_p2c, 4), "GBps": round(10 * s * n_arr / ms_p2c / 1e6, 1)}
        extra["time_step_f64"] = {"ms": round(ms_ts, 4), "GBps": round(5 * s * n_arr / ms_ts / 1e6, 1)}
        # Fused full step (godunov + cons_to_prim + next dt, dt on device).
        sim = eb.simulation.Simulation(ex, prim, cons, eos, mesh, 0.5)
        sim.run(4, use_graph=True)  # eager first pair, then graph capture + one replay
        ms_fused = timer.run(lambda: sim.run(2, use_graph=True), max(2, reps // 2)) / 2
        fused_bytes = s * (5 * n_arr + 15 * n_int)
        extra["step_fused_f64"] = {
            "ms": round(ms_fused, 4),
            "Mcell_updates_per_s": round(n_int / ms_fused / 1e3, 1),
            "GBps_algorithmic": round(fused_bytes / ms_fused / 1e6, 1),
            "hbm_frac": round(fused_bytes / ms_fused / 1e6 / HBM_PEAK_GBS, 4),
        }
        # Reference-order full step: time_step (sync) + godunov + cons_to_prim.
        def ref_step():
            d = eb.time_step(ex, prim, eos, mesh)
            eb.godunov(ex, prim, cons, eos, mesh, hllc, 0.5 * d)
            eb.cons_to_prim(ex, cons, prim, eos)
        ms_refstep = timer.run(ref_step, max(2, reps // 2))
        extra["step_reference_order_f64_ms"] = round(ms_refstep, 4)
        del sim
        log("FP32 state")
        del prim, cons
        torch.cuda.empty_cache()
        prim32, cons32 = alloc(torch.float32)
        dt32 = 0.5 * eb.time_step(ex, prim32, eos, mesh)
        timer.run(lambda: eb.godunov(ex, prim32, cons32, eos, mesh, hllc, dt32), 2)
        ms32 = timer.run(lambda: eb.godunov(ex, prim32, cons32, eos, mesh, hllc, dt32), args.steps)
        extra["godunov_f32"] = {
            "ms": round(ms32, 4),
            "Mcell_updates_per_s": round(n_int / ms32 / 1e3, 1),
            "GBps_algorithmic": round(4 * (5 * n_arr + 10 * n_int) / ms32 / 1e6, 1),
            "speedup_vs_f64": round(ms / ms32, 3),
        }
        ms_sp32, sp32_gbs = stream_probe(torch, eb, ex, timer, n_arr, torch.float32)
        g32 = 4 * (5 * n_arr + 10 * n_int) / ms32 / 1e6
        extra["stream_probe_godunov_traffic_f32"] = {
            "kernel": "eb_stream_probe_f32: b.f += a.f for 5 fields over 514^3 FP32 arrays (4-byte accesses)",
            "ms": round(ms_sp32, 4), "GBps": round(sp32_gbs, 1), "hbm_frac": round(sp32_gbs / HBM_PEAK_GBS, 4),
            "godunov_f32_frac_of_probe": round(g32 / sp32_gbs, 4),
            "probe_speedup_f32_vs_f64": round(ms_sprobe / ms_sp32, 3)}
        # Fused full step in FP32 on the same state (graph-replayed pairs).
        sim32 = eb.simulation.Simulation(ex, prim32, cons32, eos, mesh, 0.5)
        sim32.run(4, use_graph=True)
        ms_fused32 = timer.run(lambda: sim32.run(2, use_graph=True), max(2, reps // 2)) / 2
        extra["step_fused_f32"] = {
            "ms": round(ms_fused32, 4),
            "Mcell_updates_per_s": round(n_int / ms_fused32 / 1e3, 1),
            "GBps_algorithmic": round(4 * (5 * n_arr + 15 * n_int) / ms_fused32 / 1e6, 1),
            "speedup_vs_f64": round(ms_fused / ms_fused32, 3),
        }
        del sim32, prim32, cons32
        torch.cuda.empty_cache()

    if not args.quick and not args.no_fp32_report:
        log("FP32 vs FP64 implosion report")
        extra["fp32_vs_fp64"] = fp32_report(torch, eb, ex, n=n)
        log(f"fp32 report: {extra['fp32_vs_fp64']}")
        torch.cuda.empty_cache()

    if not args.quick and not args.no_batch:
        log("batched ensemble: 256 FP32 implosions of uneven shapes")
        extra["batch"] = batch_leg(torch, eb, ex, timer)
        log(f"batch: {extra['batch']}")
        torch.cuda.empty_cache()

    if not args.quick and not args.no_adjoint:
        log("adjoint: 128^3 implosion, 200 steps, forward + adjoint and FD check")
        extra["adjoint"] = adjoint_leg(torch, eb, ex, timer)
        log(f"adjoint: {extra['adjoint']}")
        torch.cuda.empty_cache()

    if args.copy_calibration:
        src = torch.empty(1 << 29, dtype=torch.float64, device=dev)  # 4 GiB
        dst = torch.empty_like(src)
        src.fill_(1.0)
        ms_copy = timer.run(lambda: dst.copy_(src), 5)
        extra["copy_calibration"] = {"bytes_read": src.numel() * 8, "bytes_written": dst.numel() * 8,
                                     "ms": round(ms_copy, 4), "GBps": round(2 * src.numel() * 8 / ms_copy / 1e6, 1)}
        del src, dst

    cpu = None
    if not args.no_cpu:
        log("CPU oracle baseline")
        cpu = cpu_baseline(n)
        if cpu is not None:
            log(f"CPU: {cpu['value']:.2f} Mcell-updates/s on {cpu['cores']} threads ({cpu['cpu_model']})")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"godunov step (HLLC, perfect gas), {n}^3 interior cells in {na}^3 arrays, FP64, "
                        f"int64 indexing, seeded random state rho~U[0.5,2] u~N(0,0.3) p~U[0.5,2] seed 1234, "
                        f"gamma 1.4, dt = 0.5*time_step",
            "interior_cells": n_int,
            "array_cells": n_arr,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(kernel_name(eb), n),
            "kernel": kernel_name(eb),
            "bytes_per_launch": bytes_godunov,
            "avg_launch_ms": round(ms, 4),
            # the same traffic per cell streamed with no arithmetic, tiles or barriers
            # (extra.stream_probe_godunov_traffic; not run with --quick)
            "stream_probe_GBps": extra.get("stream_probe_godunov_traffic", {}).get("GBps"),
        },
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)
    ex.close()


if __name__ == "__main__":
    main()
