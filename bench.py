#!/usr/bin/env python3
"""Benchmark of the master-equation RK4 hot path on one MI355X.

Workload (n=28 precision configuration): 2^28 occupation states, alpha and
beta ~ U[0.2, 1.0], 27 bulk rates ~ LogNormal(0, 0.3) (seed 28), uniform
initial state, h = 0.01, classical RK4 with an FP32 state vector and FP32
generator arithmetic (FP64 checksum, mass and moments; the moments are
divided by the FP64 mass), matrix-free generator.  A "step" is
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
    # the same loop on one thread, on a smaller lattice (a step at n=28 takes
    # about a minute there)
    n1 = min(n, 22)
    d1 = 1 << n1
    rp1, ci1, va1 = oracle.gencrs(n1, workload_rates(n1))
    x1v = oracle.uniform_x0(n1)
    k1 = [np.zeros(d1) for _ in range(4)]
    xx1 = np.zeros(d1)
    gomp.omp_set_num_threads(1)
    try:
        t1 = time.perf_counter()
        s1 = 0
        while True:
            for i, k in enumerate(k1):
                if use_ref:
                    R.rk4_step(i + 1, d1, k, x1v, xx1, rp1, ci1, va1, h)
                else:
                    L.tor_apply_csr(d1, rp1, ci1, va1, x1v, h, k, 1)
            L.tor_ref_update4(d1, x1v, k1[0], k1[1], k1[2], k1[3])
            s1 += 1
            e1 = time.perf_counter() - t1
            if e1 > budget_s / 4 and s1 >= 2:
                break
    finally:
        gomp.omp_set_num_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(value=rate, unit="state-updates/s", cores=threads,
                kind="reference" if use_ref else "port",
                sample=(f"n={n}, {steps} steps of the -f CSR loop (4 row-parallel CSR "
                        f"SpMVs {'by the reference rk4_step built from source' if use_ref else 'restated'}"
                        f" + serial update), FP64, {threads} OpenMP threads, {el:.1f} s"),
                one_thread={"value": s1 * d1 / e1, "n": n1, "steps": s1, "seconds": e1},
                cpu_model=model)


def step_h(T_end, t0, steps):
    """h = (T - t)/(long double)steps, as the reference driver computes it."""
    return float(np.longdouble(float(T_end) - float(t0)) / np.longdouble(steps))


def timed_run(ev, e, t0, h, steps, **kw):
    a, b = ev.event(), ev.event()
    ev.device_sync()
    ev.record(a, e.stream)
    info = e.run(t0, h, steps, **kw)
    ev.record(b, e.stream)
    return info, ev.elapsed_ms(a, b)


def config_1(T, ev):
    """n=10, alpha=beta=1/2, gamma=1, empty lattice, t=0..100, 10000 steps,
    reference loop in FP64, CSR and matrix-free."""
    n = 10
    r = np.array([0.5] + [1.0] * (n - 1) + [0.5])
    h = step_h(100.0, 0.0, 10000)
    res, xs = {}, {}
    # mf: the LDS-resident whole-run kernel (default for n <= 13);
    # mf_streaming: one launch per step over HBM
    for name, gen, lds in (("csr", "csr", 1), ("mf", "mf", 1), ("mf_streaming", "mf", 0)):
        with T.Engine(n, "f64", "ref", gen) as e:
            e.set_option(T.OPT_LDS, lds)
            e.set_rates(r)
            e.run(0.0, h, 20)                  # warm-up: code objects, CSR build
            e.set_x0_state(0)
            info, ms = timed_run(ev, e, 0.0, h, 10000)
            xs[name] = e.get_x()
            y, _, _, mass = e.moments()
        res[name] = {"ms": ms, "steps_done": info.steps_done, "stopped": info.stopped,
                     "state_updates_per_s": info.steps_done * (1 << n) / (ms * 1e-3),
                     "t": info.t, "mass": mass, "density_entry_to_exit": list(y[::-1])}
    res["csr_equals_mf_bitwise"] = all(
        bool(np.array_equal(xs["csr"].view(np.uint64), xs[k].view(np.uint64)))
        for k in ("mf", "mf_streaming"))
    return res


def config_2(T, ev):
    """n=20, rates ~ U[0.1,2] (seed 20), uniform start, t=0..200, 20000 steps,
    early exit; CSR vs matrix-free (reference loop) and RK4 FP64."""
    n = 20
    r = np.random.default_rng(20).uniform(0.1, 2.0, n + 1)
    h = step_h(200.0, 0.0, 20000)
    res, xs = {}, {}
    for name, integ, gen, thr in (("ref_csr", "ref", "csr", 1e-7), ("ref_mf", "ref", "mf", 1e-7),
                                  ("rk4_mf", "rk4", "mf", 1e-7), ("rk4_mf_thr1e-10", "rk4", "mf", 1e-10)):
        with T.Engine(n, "f64", integ, gen) as e:
            e.set_rates(r)
            e.run(0.0, h, 20)                  # warm-up: code objects, CSR build
            e.set_x0_uniform()
            info, ms = timed_run(ev, e, 0.0, h, 20000, thr=thr, patience=100)
            if name.startswith("ref"):
                xs[name] = e.get_x()
            y, _, _, mass = e.moments()
        res[name] = {"ms": ms, "steps_done": info.steps_done, "stopped": info.stopped,
                     "steps_at_stop": info.steps_at_stop,
                     "state_updates_per_s": info.steps_done * (1 << n) / (ms * 1e-3),
                     "t": info.t, "mass": mass}
    res["ref_csr_equals_mf_bitwise"] = bool(np.array_equal(xs["ref_csr"].view(np.uint64),
                                                            xs["ref_mf"].view(np.uint64)))
    return res


def config_3(T, ev):
    """n=28 precision: FP32 state vs FP64, 5000 steps (t=0..50)."""
    n, h, steps = 28, 0.01, 5000
    r = workload_rates(n)
    out = {}
    mom = {}
    for prec in ("f32", "f64"):
        with T.Engine(n, prec, "rk4", "mf") as e:
            e.set_rates(r)
            e.set_x0_uniform()
            info, ms = timed_run(ev, e, 0.0, h, steps, patience=-1)
            y, cov, _, mass = e.moments()
        mom[prec] = (y, cov)
        out[prec] = {"ms": ms, "ms_per_step": ms / steps,
                     "state_updates_per_s": steps * (1 << n) / (ms * 1e-3),
                     "mass_before_normalisation": info.mass, "t": info.t}
    out["max_density_err_f32_vs_f64"] = float(np.max(np.abs(mom["f32"][0] - mom["f64"][0])))
    out["max_cov_err_f32_vs_f64"] = float(np.max(np.abs(mom["f32"][1] - mom["f64"][1])))
    out["speedup_f32_over_f64"] = out["f32"]["state_updates_per_s"] / out["f64"]["state_updates_per_s"]
    # the reference's own loop (FP64, bit-faithful, first-order in h) on the
    # same workload, for the drop-in path's throughput
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        info, ms = timed_run(ev, e, 0.0, h, steps, patience=-1)
        y, _, _, _ = e.moments()
    out["ref_f64"] = {"ms": ms, "ms_per_step": ms / steps,
                      "state_updates_per_s": steps * (1 << n) / (ms * 1e-3),
                      "algorithmic_GBps": steps * (1 << n) * 16 / (ms * 1e-3) / 1e9,
                      "max_density_diff_vs_rk4_f64": float(np.max(np.abs(y - mom["f64"][0])))}
    return out


def config_5(T, ev):
    """n=32: alpha=0.3, beta=0.7, gamma=1 from the empty lattice, h=0.01;
    FP32 1000 steps (matrix-free), drift against an FP64 run at step 100."""
    n, h = 32, 0.01
    r = np.array([0.3] + [1.0] * (n - 1) + [0.7])
    out = {}
    with T.Engine(n, "f32", "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_state(0)
        info, ms100 = timed_run(ev, e, 0.0, h, 100, patience=-1)
        y32_100, _, _, _ = e.moments()
        info, ms = timed_run(ev, e, 0.0, h, 900, patience=-1, resume=True)
        y32, cov32, _, mass = e.moments()
        out["f32"] = {"ms_1000_steps": ms100 + ms, "ms_per_step": (ms100 + ms) / 1000,
                      "state_updates_per_s": 1000 * (1 << n) / ((ms100 + ms) * 1e-3),
                      "t": info.t, "steps": info.steps_done, "mass": mass,
                      "density_entry_to_exit_t10": list(y32[::-1])}
    with T.Engine(n, "f64", "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_state(0)
        info, ms = timed_run(ev, e, 0.0, h, 100, patience=-1)
        y64_100, _, _, _ = e.moments()
        out["f64"] = {"ms_100_steps": ms, "ms_per_step": ms / 100,
                      "state_updates_per_s": 100 * (1 << n) / (ms * 1e-3)}
    out["max_density_drift_f32_vs_f64_step100"] = float(np.max(np.abs(y32_100 - y64_100)))
    return out


def config_4_instances(count=512, seed=512):
    """512 instances: n ~ U{8..18}, n+1 rates ~ LogNormal(0, 0.5); odd
    instances time-variant r_i(t) = A_i sin(w_i t) + 1, A ~ U[0, 0.5],
    w ~ U[0.5, 2]."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(8, 19, size=count)
    spec = []
    for i, n in enumerate(ns):
        n = int(n)
        r = rng.lognormal(0.0, 0.5, n + 1)
        if i % 2 == 1:
            spec.append(("tv", dict(type=np.zeros(n + 1, np.int32), amp=rng.uniform(0, 0.5, n + 1),
                                    period=rng.uniform(0.5, 2.0, n + 1), offset=np.zeros(n + 1),
                                    c=np.ones(n + 1))))
        else:
            spec.append(("r", r))
    return [int(v) for v in ns], spec


def config_4(T, ev, steps=2000, h=0.01, sequential=True):
    """Batching: 512 instances in one batch (LDS-resident small instances in
    one launch, the rest sharing one launch per RK4 pass, graph-replayed)
    vs a sequential loop of single-instance runs; FP32, uniform start."""
    ns, spec = config_4_instances()
    total_states = sum(1 << n for n in ns)
    out = {"instances": len(ns), "total_states": total_states, "steps": steps,
           "n_histogram": {str(n): ns.count(n) for n in sorted(set(ns))}}
    with T.Batch(ns, "f32") as b:
        for i, (kind, val) in enumerate(spec):
            if kind == "tv":
                b.set_time_variant(i, val)
            else:
                b.set_rates(i, val)
        b.run(0.0, h, 64)                      # builds tables and the graph
        b.set_x0_uniform()
        a, e = ev.event(), ev.event()
        ev.device_sync()
        ev.record(a, b.stream)
        infos = b.run(0.0, h, steps)
        ev.record(e, b.stream)
        ms = ev.elapsed_ms(a, e)
        xb = {i: b.get_x(i) for i in range(0, len(ns), 37)}
    out["batched"] = {"ms": ms, "state_updates_per_s": total_states * steps / (ms * 1e-3),
                      "all_steps_done": all(inf.steps_done == steps for inf in infos),
                      "max_mass_error": max(abs(inf.mass - 1.0) for inf in infos)}
    # algorithmic HBM bytes of the streamed instances (n >= 14 in FP32):
    # Horner 44 B, stored stages 68 B per state and step
    streamed = [(n, k) for n, (k, _) in zip(ns, spec) if n >= 14]
    algo = sum((1 << n) * (44 if k == "r" else 68) for n, k in streamed) * steps
    out["batched"]["streamed_algorithmic_GBps"] = algo / (ms * 1e-3) / 1e9
    if sequential:
        seq_ms = 0.0
        maxdiff = 0.0
        for i, (n, (kind, val)) in enumerate(zip(ns, spec)):
            with T.Engine(n, "f32", "rk4", "mf") as eng:
                if kind == "tv":
                    eng.set_time_variant(val)
                else:
                    eng.set_rates(val)
                eng.set_x0_uniform()
                _, m = timed_run(ev, eng, 0.0, h, steps, patience=-1)
                seq_ms += m
                if i in xb:
                    maxdiff = max(maxdiff, float(np.max(np.abs(eng.get_x() - xb[i]))))
        out["sequential"] = {"ms": seq_ms, "state_updates_per_s": total_states * steps / (seq_ms * 1e-3),
                             "max_abs_diff_vs_batched_sampled": maxdiff}
        out["speedup_batched_vs_sequential"] = seq_ms / ms
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=28)
    ap.add_argument("--h", type=float, default=0.01)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--config", type=int, default=0,
                    help="run one of the five benchmark workloads (1..5) at full size instead")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(see the sharding design); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import tasep_amd as T
    T.load_library()
    ev = HipEvents()
    if args.config:
        fn = {1: config_1, 2: config_2, 3: config_3, 4: config_4, 5: config_5}.get(args.config)
        if fn is None:
            sys.exit(f"workload {args.config} is not available")
        print(json.dumps({"config": args.config, "result": fn(T, ev)}), flush=True)
        return
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
                   "rk4_form": "horner", "storage": "f32",
                   "arithmetic": "f32 generator, f64 checksum/mass/moments",
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
