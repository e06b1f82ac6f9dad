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
steps and after the whole 5000-step horizon, the roofline of the dominant
kernel (per-launch HIP-event timing on the engine's stream), a CPU baseline
(the reference's CSR path -- its own rk4_step compiled from source when
oracle/_ref exists, else the oracle's restatement -- timed on a bounded
sample of the same n on this host) and the FP64 GPU results against FP64
CPU results on the same workload: the drop-in reference loop against that
CPU loop's state bit for bit, RK4 FP64 against the oracle's RK4 FP64.
Progress lines go to stderr.

    python bench.py --config K      one of the five workloads at full size,
                                    each with its roofline and CPU baseline
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


def progress(msg):
    """One line per finished phase on stderr (stdout carries only the JSON)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


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


def roofline_entry(e, n, h, reps=3):
    """The roofline object of one engine's dominant kernel (per-launch HIP
    events on the engine's stream; algorithmic bytes per launch)."""
    name, avg_ms, algo_bytes, achieved, prof = roofline(e, n, h, reps)
    return {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "algo_bytes_per_launch": algo_bytes,
            "avg_launch_ms": avg_ms,
            "step_ms": sum(m for _, m, _ in prof),
            "kernels": [{"name": a, "ms": b, "bytes_per_state": c} for a, b, c in prof]}


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


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def _omp():
    return C.CDLL("libgomp.so.1")


class CpuReferenceLoop:
    """The reference's fastest CPU path at size n: its -f CSR loop, four
    row-parallel rk4_step SpMVs per step (the reference's own solv.c built
    from source into oracle/_ref when present, else the oracle's
    restatement) and the serial update, from the uniform state."""

    def __init__(self, n, rates, h):
        import oracle
        self.oracle = oracle
        self.n, self.h = n, h
        self.dim = 1 << n
        self.rowPtr, self.colInd, self.val = oracle.gencrs(n, rates)
        self.x = oracle.uniform_x0(n)
        self.ks = [np.zeros(self.dim) for _ in range(4)]
        self.x1 = np.zeros(self.dim)
        self.use_ref = oracle.have_ref()
        self.L = oracle.lib()
        self.R = oracle.ref_lib() if self.use_ref else None
        self.steps = 0

    def step(self, threads):
        for i, k in enumerate(self.ks):
            if self.use_ref:
                self.R.rk4_step(i + 1, self.dim, k, self.x, self.x1, self.rowPtr, self.colInd, self.val,
                                self.h)
            else:
                self.L.tor_apply_csr(self.dim, self.rowPtr, self.colInd, self.val, self.x, self.h, k, threads)
        self.L.tor_ref_update4(self.dim, self.x, self.ks[0], self.ks[1], self.ks[2], self.ks[3])
        self.steps += 1

    def timed(self, budget_s, threads, min_steps=2):
        """Steps until budget_s has passed (at least min_steps); returns
        (steps, seconds) of the timed part."""
        gomp = _omp()
        keep = int(gomp.omp_get_max_threads())
        gomp.omp_set_num_threads(threads)
        try:
            t0 = time.perf_counter()
            s = 0
            while True:
                self.step(threads)
                s += 1
                el = time.perf_counter() - t0
                if el > budget_s and s >= min_steps:
                    break
        finally:
            gomp.omp_set_num_threads(keep)
        return s, el


def _fits_host(n):
    """Whether the CPU CSR loop at size n fits in 60% of the free host memory."""
    need = ((1 << n) + 1) * 8 + _nnz(n) * 16 + 6 * (1 << n) * 8
    try:
        avail = os.sysconf("SC_AVPHYS_PAGES") * os.sysconf("SC_PAGE_SIZE")
    except (ValueError, OSError):
        return True
    return need <= 0.6 * avail


def _nnz(n):
    return 2 * (1 << n) + (n - 1) * (1 << (n - 2))


def cpu_baseline(n, rates_fn, h, budget_s=20.0, keep_loop=False):
    """cpu_baseline object: the reference's CSR loop timed on a bounded
    sample at size n (the largest size <= n whose matrix fits the host) with
    every OpenMP thread, and the same loop on one thread at min(n, 22).
    rates_fn(m) gives the rates at size m.  With keep_loop the multi-thread
    loop object (its state after loop.steps steps) is returned too."""
    threads = int(_omp().omp_get_max_threads())
    m = n
    while m > 20 and not _fits_host(m):
        m -= 2
    loop = CpuReferenceLoop(m, rates_fn(m), h)
    progress(f"CPU CSR matrix built at n={m}")
    loop.step(threads)                        # first touch of every array
    steps, el = loop.timed(budget_s, threads)
    n1 = min(m, 22)
    one = CpuReferenceLoop(n1, rates_fn(n1), h)
    s1, e1 = one.timed(budget_s / 4, 1)
    out = dict(value=steps * loop.dim / el, unit="state-updates/s", cores=threads,
               kind="reference" if loop.use_ref else "port",
               sample=(f"n={m}, {steps} steps of the -f CSR loop (4 row-parallel CSR SpMVs "
                       f"{'by the reference rk4_step built from source' if loop.use_ref else 'restated'}"
                       f" + serial update), FP64, {threads} OpenMP threads, {el:.1f} s"),
               one_thread={"value": s1 * one.dim / e1, "n": n1, "steps": s1, "seconds": e1},
               cpu_model=_cpu_model())
    if keep_loop:
        return out, loop
    del loop
    return out


def ref_loop_vs_cpu(T, n, rates, h, steps, x0_state=None, gen="mf"):
    """The drop-in reference loop on the GPU against the reference's own CSR
    loop on the host (CpuReferenceLoop), `steps` steps from the uniform
    state or from the basis state x0_state: bit-exactness and the largest
    differences."""
    import oracle
    loop = CpuReferenceLoop(n, rates, h)
    if x0_state is not None:
        loop.x[:] = 0.0
        loop.x[x0_state] = 1.0
    threads = int(_omp().omp_get_max_threads())
    for _ in range(steps):
        loop.step(threads)
    with T.Engine(n, "f64", "ref", gen) as e:
        e.set_rates(rates)
        if x0_state is None:
            e.set_x0_uniform()
        else:
            e.set_x0_state(x0_state)
        e.run(0.0, h, steps, patience=-1)
        xg = e.get_x()
    return {"n": n, "steps": steps, "generator": gen,
            "bit_exact": bool(np.array_equal(xg.view(np.uint64), loop.x.view(np.uint64))),
            "max_abs_x_diff": float(np.max(np.abs(xg - loop.x))),
            "max_density_diff": float(np.max(np.abs(oracle.site_moments(n, xg) - oracle.site_moments(n, loop.x))))}


def fp64_gpu_vs_cpu(T, loop, rates, h, rk4_steps=2):
    """FP64 GPU against FP64 CPU on the same workload: the drop-in reference
    loop (matrix-free tiles) over the CPU reference loop's steps, compared
    bit for bit, and classical RK4 FP64 on the GPU against the oracle's FP64
    RK4 on the host (a few steps), compared on every state and density."""
    import oracle
    n = loop.n
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_rates(rates)
        e.set_x0_uniform()
        e.run(0.0, h, loop.steps, patience=-1)
        xg = e.get_x()
    ref = {"steps": loop.steps, "bit_exact": bool(np.array_equal(xg.view(np.uint64), loop.x.view(np.uint64))),
           "max_abs_x_diff": float(np.max(np.abs(xg - loop.x))),
           "max_density_diff": float(np.max(np.abs(oracle.site_moments(n, xg) - oracle.site_moments(n, loop.x))))}
    del xg
    x0 = oracle.uniform_x0(n)
    t0 = time.perf_counter()
    xc, _ = oracle.run_rk4(n, rates, x0, 0.0, h, rk4_steps)
    cpu_s = time.perf_counter() - t0
    with T.Engine(n, "f64", "rk4", "mf") as e:
        e.set_rates(rates)
        e.set_x0_uniform()
        e.run(0.0, h, rk4_steps, patience=-1)
        xg = e.get_x()
    # both densities from the same sequential host sums, so the difference
    # is the states' own
    rk4 = {"steps": rk4_steps, "max_rel_x_diff": float(np.max(np.abs(xg - xc) / np.abs(xc))),
           "max_density_diff": float(np.max(np.abs(oracle.site_moments(n, xg) - oracle.site_moments(n, xc)))),
           "cpu_seconds": cpu_s}
    return {"n": n, "ref_loop_vs_reference_cpu_loop": ref, "rk4_f64_vs_oracle_rk4_f64": rk4}


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
        progress(f"config 1 {name}: {ms:.1f} ms")
    res["csr_equals_mf_bitwise"] = all(
        bool(np.array_equal(xs["csr"].view(np.uint64), xs[k].view(np.uint64)))
        for k in ("mf", "mf_streaming"))
    # one launch runs all 10,000 steps with both state copies in LDS: HBM is
    # touched once, so the fraction below (16 B per state-update over the
    # run) only says how far from a streaming kernel's bound it runs
    mf = res["mf"]
    ach = mf["state_updates_per_s"] * 16 / 1e9
    res["roofline"] = {"bound": "latency", "kernel": "k_lds_run (whole run in LDS)", "achieved": ach,
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                       "algo_bytes_per_state_update": 16}
    res["cpu_baseline"] = cpu_baseline(n, lambda m: np.array([0.5] + [1.0] * (m - 1) + [0.5]), h, budget_s=5.0)
    res["speedup_vs_cpu"] = mf["state_updates_per_s"] / res["cpu_baseline"]["value"]
    # the whole run against the reference's own loop on the host
    res["fp64_gpu_vs_cpu"] = {g: ref_loop_vs_cpu(T, n, r, h, 10000, x0_state=0, gen=g) for g in ("mf", "csr")}
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
            roof = roofline_entry(e, n, h)
        res[name] = {"ms": ms, "steps_done": info.steps_done, "stopped": info.stopped,
                     "steps_at_stop": info.steps_at_stop,
                     "state_updates_per_s": info.steps_done * (1 << n) / (ms * 1e-3),
                     "t": info.t, "mass": mass, "roofline": roof}
        progress(f"config 2 {name}: {ms:.1f} ms")
    res["ref_csr_equals_mf_bitwise"] = bool(np.array_equal(xs["ref_csr"].view(np.uint64),
                                                            xs["ref_mf"].view(np.uint64)))
    res["cpu_baseline"] = cpu_baseline(n, lambda m: np.random.default_rng(20).uniform(0.1, 2.0, m + 1), h,
                                       budget_s=15.0)
    res["speedup_ref_mf_vs_cpu"] = res["ref_mf"]["state_updates_per_s"] / res["cpu_baseline"]["value"]
    # the first 200 steps against the reference's own loop on the host
    res["fp64_gpu_vs_cpu"] = {g: ref_loop_vs_cpu(T, n, r, h, 200, gen=g) for g in ("mf", "csr")}
    progress("config 2 FP64 GPU vs CPU compared")
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
            roof = roofline_entry(e, n, h)
        mom[prec] = (y, cov)
        out[prec] = {"ms": ms, "ms_per_step": ms / steps,
                     "state_updates_per_s": steps * (1 << n) / (ms * 1e-3),
                     "mass_before_normalisation": info.mass, "t": info.t, "roofline": roof}
        progress(f"config 3 {prec}: {ms / steps:.3f} ms/step")
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
        roof = roofline_entry(e, n, h)
    out["ref_f64"] = {"ms": ms, "ms_per_step": ms / steps,
                      "state_updates_per_s": steps * (1 << n) / (ms * 1e-3),
                      "algorithmic_GBps": steps * (1 << n) * 16 / (ms * 1e-3) / 1e9,
                      "max_density_diff_vs_rk4_f64": float(np.max(np.abs(y - mom["f64"][0]))),
                      "roofline": roof}
    progress(f"config 3 reference loop: {ms / steps:.3f} ms/step")
    cpu, loop = cpu_baseline(n, workload_rates, h, keep_loop=True)
    out["cpu_baseline"] = cpu
    progress("config 3 CPU baseline done")
    out["fp64_gpu_vs_cpu"] = fp64_gpu_vs_cpu(T, loop, workload_rates(loop.n), h)
    del loop
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
                      "density_entry_to_exit_t10": list(y32[::-1]),
                      "roofline": roofline_entry(e, n, h, reps=2)}
    progress("config 5 FP32 done")
    with T.Engine(n, "f64", "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_state(0)
        info, ms = timed_run(ev, e, 0.0, h, 100, patience=-1)
        y64_100, _, _, _ = e.moments()
        out["f64"] = {"ms_100_steps": ms, "ms_per_step": ms / 100,
                      "state_updates_per_s": 100 * (1 << n) / (ms * 1e-3),
                      "roofline": roofline_entry(e, n, h, reps=2)}
    out["max_density_drift_f32_vs_f64_step100"] = float(np.max(np.abs(y32_100 - y64_100)))
    progress("config 5 FP64 done")
    # the reference cannot run n=32 (its unsigned int state loops never end
    # at 2^32 states, and the -f matrix would need 670 GB): its CSR loop is
    # timed at the largest lattice whose matrix fits this host, same rates
    out["cpu_baseline"] = cpu_baseline(28, lambda m: np.array([0.3] + [1.0] * (m - 1) + [0.7]), h)
    out["cpu_baseline"]["note"] = "n=32 is beyond the reference; its loop timed at the size in `sample`"
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
    progress(f"config 4 batch: {ms:.1f} ms")
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
        progress(f"config 4 sequential: {seq_ms:.1f} ms")
    out["errors"] = batch_errors(T, ns, spec, h, steps, xb)
    progress("config 4 errors against FP64 done")
    ach = out["batched"]["streamed_algorithmic_GBps"]
    out["roofline"] = {"bound": "hbm", "kernel": "k_batch_bigtile / k_batch_tile / k_batch_pass (run average, streamed instances)",
                       "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS}
    out["cpu_baseline"] = cpu_batch_baseline(ns, spec, h)
    out["speedup_batched_vs_cpu"] = out["batched"]["state_updates_per_s"] / out["cpu_baseline"]["value"]
    return out


def batch_errors(T, ns, spec, h, steps, xb):
    """FP32 batch against the same batch in FP64 (every instance's densities)
    and the sampled FP32 states against the FP64 oracle's RK4 on the host."""
    import oracle
    with T.Batch(ns, "f64") as b:
        for i, (kind, val) in enumerate(spec):
            if kind == "tv":
                b.set_time_variant(i, val)
            else:
                b.set_rates(i, val)
        b.run(0.0, h, steps)
        y64 = [b.moments(i)[0] for i in range(len(ns))]
    with T.Batch(ns, "f32") as b:
        for i, (kind, val) in enumerate(spec):
            if kind == "tv":
                b.set_time_variant(i, val)
            else:
                b.set_rates(i, val)
        b.run(0.0, h, steps)
        y32 = [b.moments(i)[0] for i in range(len(ns))]
    dens = max(float(np.max(np.abs(a - c))) for a, c in zip(y32, y64))
    worst = 0.0
    picked = []
    for i, x in sorted(xb.items()):
        n = ns[i]
        if n > 16:
            continue
        kind, val = spec[i]
        x0 = oracle.uniform_x0(n)
        ox, _ = (oracle.run_rk4(n, None, x0, 0.0, h, steps, tv=val) if kind == "tv"
                 else oracle.run_rk4(n, val, x0, 0.0, h, steps))
        worst = max(worst, float(np.max(np.abs(x - ox))))
        picked.append(n)
    return {"max_density_err_f32_vs_f64_batch_all_instances": dens,
            "max_abs_x_err_f32_vs_fp64_oracle_sampled": worst, "oracle_sampled_n": picked}


def cpu_batch_baseline(ns, spec, h, every=37, budget_s=1.5):
    """The reference's CSR loop (one process per instance in the reference)
    on every `every`-th instance of the batch, each for about budget_s with
    every OpenMP thread; time-variant instances with their rates at t = 0
    (the reference's time-variant loop is its serial matrix-free path)."""
    threads = int(_omp().omp_get_max_threads())
    states = secs = 0.0
    picked = []
    for i in range(0, len(ns), every):
        kind, val = spec[i]
        r = val if kind == "r" else np.maximum(val["amp"] * np.sin(val["offset"]) + val["c"], 0.0)
        loop = CpuReferenceLoop(ns[i], r, h)
        loop.step(threads)
        st, el = loop.timed(budget_s, threads)
        states += st * loop.dim
        secs += el
        picked.append(ns[i])
    return dict(value=states / secs, unit="state-updates/s", cores=threads,
                kind="reference" if _have_ref() else "port",
                sample=(f"{len(picked)} of the batch's instances (every {every}th, n = {picked}), the -f CSR "
                        f"loop of each for ~{budget_s} s, FP64, {threads} OpenMP threads, {secs:.1f} s in all"),
                cpu_model=_cpu_model())


def _have_ref():
    import oracle
    return oracle.have_ref()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=28)
    ap.add_argument("--h", type=float, default=0.01)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--horizon", type=int, default=5000,
                    help="steps of the FP32-vs-FP64 error run (the workload's full t = 0..50); 0 skips it")
    ap.add_argument("--config", type=int, default=0,
                    help="run one of the five benchmark workloads (1..5) at full size instead")
    ap.add_argument("--no-sequential", action="store_true",
                    help="workload 4: skip the sequential per-instance comparison loop")
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
        res = fn(T, ev, sequential=not args.no_sequential) if args.config == 4 else fn(T, ev)
        progress(f"configuration {args.config} done")
        print(json.dumps({"config": args.config, "result": res}), flush=True)
        return
    n, h, K, W = args.n, args.h, args.steps, args.warmup
    rates = workload_rates(n)
    dim = 1 << n

    e32, ms32, wall32, info32, y32, cov32 = time_engine(T, ev, n, "f32", rates, h, K, W)
    name, avg_ms, algo_bytes, achieved, prof = roofline(e32, n, h)
    e32.close()
    progress(f"FP32 n={n}: {ms32 / K:.3f} ms/step, dominant kernel {name} {avg_ms:.3f} ms")
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
        progress(f"FP64 n={n}: {ms64 / K:.3f} ms/step")
        v64 = K * dim / (ms64 * 1e-3)
        out["f64"] = {"value": v64, "ms_per_step": ms64 / K,
                      "speedup_f32_over_f64": value / v64,
                      "max_density_err_f32_vs_f64": float(np.max(np.abs(y32 - y64))),
                      "max_cov_err_f32_vs_f64": float(np.max(np.abs(cov32 - cov64))),
                      "t_end": info64.t}
    if not args.no_f64 and args.horizon > 0:
        out["horizon"] = horizon_error(T, n, rates, h, args.horizon)
        progress(f"horizon {args.horizon} steps: max density error {out['horizon']['max_density_err_f32_vs_f64']:.2e}")
    if not args.no_cpu:
        cpu, loop = cpu_baseline(n, workload_rates, h, keep_loop=True)
        out["cpu_baseline"] = cpu
        progress(f"CPU baseline: {cpu['value']:.3e} state-updates/s ({cpu['sample']})")
        out["fp64_gpu_vs_cpu"] = fp64_gpu_vs_cpu(T, loop, workload_rates(loop.n), h)
        progress("FP64 GPU vs CPU compared")
        del loop
    print(json.dumps(out), flush=True)


def horizon_error(T, n, rates, h, steps):
    """FP32 against FP64 after the workload's whole horizon (5000 steps,
    t = 0..50) from the uniform state: max per-site density and covariance
    errors (moments of the FP32 state divided by its FP64 mass)."""
    mom = {}
    for prec in ("f32", "f64"):
        with T.Engine(n, prec, "rk4", "mf") as e:
            e.set_rates(rates)
            e.set_x0_uniform()
            info = e.run(0.0, h, steps, patience=-1)
            y, cov, _, mass = e.moments()
        mom[prec] = (y, cov, info.t)
        progress(f"horizon run {prec} done")
    return {"steps": steps, "t_end": mom["f64"][2],
            "max_density_err_f32_vs_f64": float(np.max(np.abs(mom["f32"][0] - mom["f64"][0]))),
            "max_cov_err_f32_vs_f64": float(np.max(np.abs(mom["f32"][1] - mom["f64"][1])))}


if __name__ == "__main__":
    main()
