#!/usr/bin/env python3
"""Benchmark of the master-equation RK4 hot path on one MI355X.

Workload (n=28 precision configuration): 2^28 occupation states, alpha and
beta ~ U[0.2, 1.0], 27 bulk rates ~ LogNormal(0, 0.3) (seed 28), uniform
initial state, h = 0.01, classical RK4 with an FP32 state vector (FP64
arithmetic, sums and renormalisation), matrix-free generator.  A "step" is
one RK4 step over all 2^28 states; the metric is state-updates/s =
steps * 2^n / s with the state resident in HBM.

    python bench.py [--steps K] [--warmup W] [--n 28] [--no-cpu] [--no-f64]

Prints one JSON line.  Beside the FP32 rate it reports the FP64 RK4 rate on
the same workload, the FP32-vs-FP64 density/covariance error after the timed
steps, the roofline of the dominant kernel (per-launch HIP-event timing on
the engine's stream) and a CPU baseline: the reference's CSR path (its own
rk4_step compiled from source when oracle/_ref exists, else the oracle's
restatement) timed on a bounded sample of the same n on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def workload_rates(n: int, seed: int = 28):
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(0.2, 1.0, 2)
    bulk = rng.lognormal(0.0, 0.3, n - 1)
    return np.concatenate([[a], bulk, [b]])


class HipEvents:
    """hipEvent timing on a given stream, via the HIP runtime already loaded
    by libtasep_hip.so."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        ev = C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(ev)) == 0
        return ev

    def record(self, ev, stream):
        assert self.hip.hipEventRecord(ev, C.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return float(ms.value)

    def device_sync(self):
        assert self.hip.hipDeviceSynchronize() == 0


def time_engine(T, ev, n, prec, rates, h, steps, warmup):
    """Run warmup steps, reset to the uniform state, time `steps` steps."""
    e = T.Engine(n, prec, "rk4", "mf")
    e.set_rates(rates)
    e.set_x0_uniform()
    if warmup:
        e.run(0.0, h, warmup, patience=-1)
    e.set_x0_uniform()
    a, b = ev.event(), ev.event()
    ev.device_sync()
    t0 = time.perf_counter()
    ev.record(a, e.stream)
    info = e.run(0.0, h, steps, patience=-1)
    ev.record(b, e.stream)
    ms = ev.elapsed_ms(a, b)
    ev.device_sync()
    wall = time.perf_counter() - t0
    assert info.steps_done == steps, info
    y, cov, _, mass = e.moments()
    return e, ms, wall, info, y, cov


def roofline(e, n, h, reps=5):
    prof = e.step_profile(0.0, h, reps=reps)
    by = {}
    for name, ms, bps in prof:
        d = by.setdefault(name, [0.0, 0, bps])
        d[0] += ms
        d[1] += 1
    # dominant kernel: largest total time per step
    name, (tot, cnt, bps) = max(by.items(), key=lambda kv: kv[1][0])
    avg_ms = tot / cnt
    algo_bytes = bps * float(1 << n)
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    return name, avg_ms, algo_bytes, achieved, prof


def measured_traffic(name, n):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary (profiles/traffic_n{n}.json), if one exists for this workload."""
    p = os.path.join(REPO, "profiles", f"traffic_n{n}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(name)
    return None if k is None else float(k["bytes_per_launch"])


def cpu_baseline(n, rates, h, budget_s=20.0):
    """The reference's fastest CPU path (-f CSR, OpenMP row-parallel SpMV x4
    per step + serial update) on a bounded number of steps at size n."""
    import oracle
    gomp = C.CDLL("libgomp.so.1")
    threads = int(gomp.omp_get_max_threads())
    dim = 1 << n
    nnz = oracle.nnz(n)
    need = (dim + 1) * 8 + nnz * 16 + 6 * dim * 8
    try:
        avail = os.sysconf("SC_AVPHYS_PAGES") * os.sysconf("SC_PAGE_SIZE")
    except (ValueError, OSError):
        avail = need * 2
    while need > 0.6 * avail and n > 20:
        n -= 2
        dim = 1 << n
        nnz = oracle.nnz(n)
        need = (dim + 1) * 8 + nnz * 16 + 6 * dim * 8
        rates = workload_rates(n)
    rowPtr, colInd, val = oracle.gencrs(n, rates)
    x = oracle.uniform_x0(n)
    ks = [np.zeros(dim) for _ in range(4)]
    x1 = np.zeros(dim)
    use_ref = oracle.have_ref()
    L = oracle.lib()
    if use_ref:
        R = oracle.ref_lib()

        def spmv(step, k):
            R.rk4_step(step, dim, k, x, x1, rowPtr, colInd, val, h)
    else:
        def spmv(step, k):
            L.tor_apply_csr(dim, rowPtr, colInd, val, x, h, k, threads)

    def one_step():
        for i, k in enumerate(ks):
            spmv(i + 1, k)
        L.tor_ref_update4(dim, x, ks[0], ks[1], ks[2], ks[3])

    one_step()  # first touch
    t0 = time.perf_counter()
    steps = 0
    while True:
        one_step()
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or (steps >= 3 and el > budget_s / 2):
            break
    rate = steps * dim / el
    return dict(value=rate, unit="state-updates/s", cores=threads,
                kind="reference" if use_ref else "port",
                sample=(f"n={n}, {steps} steps of the -f CSR loop (4 row-parallel CSR "
                        f"SpMVs {'by the reference rk4_step built from source' if use_ref else 'restated'}"
                        f" + serial update), FP64, {threads} OpenMP threads, {el:.1f} s"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=28)
    ap.add_argument("--h", type=float, default=0.01)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import tasep_amd as T
    T.load_library()
    ev = HipEvents()
    n, h, K, W = args.n, args.h, args.steps, args.warmup
    rates = workload_rates(n)
    dim = 1 << n

    e32, ms32, wall32, info32, y32, cov32 = time_engine(T, ev, n, "f32", rates, h, K, W)
    name, avg_ms, algo_bytes, achieved, prof = roofline(e32, n, h)
    e32.close()
    value = K * dim / (ms32 * 1e-3)
    out = {
        "metric": "state-updates/s (RK4 steps x 2^n / s), FP32-bulk",
        "value": value,
        "unit": "state-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms32 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded rates, uniform initial state)",
        "config": {"workload": f"n{n}_f32_rk4_mf", "n": n, "states": dim, "h": h,
                   "rk4_form": "horner", "storage": "f32", "arithmetic": "f64",
                   "rates": "alpha,beta~U[0.2,1], bulk~LogNormal(0,0.3), seed 28",
                   "wall_s": wall32},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "algo_bytes_per_launch": algo_bytes,
                     "avg_launch_ms": avg_ms,
                     "traffic": measured_traffic(name, n)},
        "step_bandwidth_GBs": sum(b for _, _, b in prof) * dim / (ms32 / K * 1e-3) / 1e9,
        "kernels": [{"name": a, "ms": b, "bytes_per_state": c} for a, b, c in prof],
    }
    if not args.no_f64:
        e64, ms64, wall64, info64, y64, cov64 = time_engine(T, ev, n, "f64", rates, h, K, W)
        e64.close()
        v64 = K * dim / (ms64 * 1e-3)
        out["f64"] = {"value": v64, "ms_per_step": ms64 / K,
                      "speedup_f32_over_f64": value / v64,
                      "max_density_err_f32_vs_f64": float(np.max(np.abs(y32 - y64))),
                      "max_cov_err_f32_vs_f64": float(np.max(np.abs(cov32 - cov64))),
                      "t_end": info64.t}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, rates, h)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
