"""CPU baseline timing of the oracle (measurement infrastructure, run by bench.py's
cpu_baseline leg as a subprocess so that the OpenMP binding variables apply).

The reference's CPU path is Kokkos' OpenMP backend (setups/adastra/genoa/job.sh:20-24
binds OMP_PROC_BIND=close, OMP_PLACES=cores); Kokkos is not available here, so the
oracle (a reference-order C++ restatement, OpenMP) stands in for it, built for
timing as the reference would be: g++ -O3 -march=native -fopenmp with the default
contraction (oracle/Makefile target `native`, compiled on the host that runs it).
The parity build (contraction off, portable ISA) is timed only when the native
build cannot be made.

Workloads, seed 1234, gamma 1.4, dx = 1/n:
  * each operator at n^3 interior cells in (n+2)^3 arrays, FP64 and FP32, on
    rho ~ U[0.5,2], u_k ~ N(0,0.3), p ~ U[0.5,2]: godunov (dt = 0.5*time_step),
    cons_to_prim, prim_to_cons, time_step; median of `reps` calls each;
  * the diagonal implosion 64^3, FP64, CFL 0.5, 100 steps (time_step + godunov +
    cons_to_prim per step), median of `reps` runs.

Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import statistics
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from oracle import Oracle  # noqa: E402


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def timed(fn, reps: int) -> float:
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def operators(orc: Oracle, n: int, dtype, reps: int) -> dict:
    na = n + 2
    gamma, dx = 1.4, 1.0 / n
    rng = np.random.default_rng(1234)
    prim = np.empty((5, na, na, na), dtype=dtype)
    prim[0] = rng.uniform(0.5, 2.0, (na, na, na))
    prim[1] = rng.uniform(0.5, 2.0, (na, na, na))
    prim[2:] = rng.normal(0.0, 0.3, (3, na, na, na))
    cons = orc.prim_to_cons(prim, gamma)
    dt = 0.5 * orc.time_step(prim, gamma, dx, dx, dx)
    s = np.dtype(dtype).itemsize
    n_int, n_arr = n**3, na**3
    out = {}
    t = timed(lambda: orc.godunov(prim, cons, gamma, dx, dx, dx, dt), reps)
    out["godunov"] = {"s": t, "Mcell_updates_per_s": n_int / t / 1e6,
                      "Mcells_per_s_ref_convention": n_arr / t / 1e6,
                      "GBps": s * (5 * n_arr + 10 * n_int) / t / 1e9}
    # conversions into preallocated outputs (first touch of a fresh array is not timed)
    work = np.empty_like(prim)
    orc.cons_to_prim(cons, gamma, out=work)
    t = timed(lambda: orc.cons_to_prim(cons, gamma, out=work), reps)
    out["cons_to_prim"] = {"s": t, "Mcells_per_s": n_arr / t / 1e6, "GBps": 10 * s * n_arr / t / 1e9}
    t = timed(lambda: orc.prim_to_cons(prim, gamma, out=work), reps)
    out["prim_to_cons"] = {"s": t, "Mcells_per_s": n_arr / t / 1e6, "GBps": 10 * s * n_arr / t / 1e9}
    t = timed(lambda: orc.time_step(prim, gamma, dx, dx, dx), reps)
    out["time_step"] = {"s": t, "Mcells_per_s": n_arr / t / 1e6, "GBps": 5 * s * n_arr / t / 1e9}
    return out


def implosion(orc: Oracle, n: int, nsteps: int, reps: int) -> dict:
    na = n + 2
    dx = 1.0 / n
    prim0 = orc.init_diagonal(na, na, na, dx, dx, dx, offset=0.45)
    cons0 = orc.prim_to_cons(prim0, 1.4)

    def run():
        p, c = prim0.copy(), cons0.copy()
        orc.simulate(p, c, 1.4, dx, dx, dx, 0.5, nsteps)

    t = timed(run, reps)
    return {"s": t, "Mcell_updates_per_s": nsteps * n**3 / t / 1e6,
            "workload": f"diagonal implosion {n}^3, FP64, CFL 0.5, {nsteps} steps"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--lib", type=str, default=None)
    ap.add_argument("--build", type=str, default="native")
    args = ap.parse_args()
    cpus = len(os.sched_getaffinity(0))  # before the OpenMP runtime binds this thread
    orc = Oracle(lib_path=Path(args.lib) if args.lib else None)
    t0 = time.perf_counter()
    res = {
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "cpus_available": cpus,
        "threads": orc.threads(),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "build": args.build,
        "n": args.n,
        "f64": operators(orc, args.n, np.float64, args.reps),
        "f32": operators(orc, args.n, np.float32, args.reps),
        "implosion_64": implosion(orc, 64, 100, args.reps),
    }
    res["seconds"] = time.perf_counter() - t0
    print(json.dumps(res))


if __name__ == "__main__":
    main()
