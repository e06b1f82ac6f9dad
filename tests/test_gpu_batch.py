"""Batched engine (tasep_batch_*) against the oracle and the single-instance
engine: every instance of a mixed batch (LDS-resident and streamed, constant
and time-variant rates) must match its own classical RK4 run."""
import numpy as np
import pytest

import oracle
import tasep_amd as T

pytestmark = pytest.mark.gpu


def _tv(n, rng):
    return dict(type=np.zeros(n + 1, np.int32), amp=rng.uniform(0, 0.5, n + 1),
                period=rng.uniform(0.5, 2, n + 1), offset=np.zeros(n + 1), c=np.ones(n + 1))


def _setup(ns, seed, prec, tv_every=2, **opts):
    rng = np.random.default_rng(seed)
    b = T.Batch(ns, prec)
    for k, v in opts.items():
        b.set_option(getattr(T, "BATCH_OPT_" + k.upper()), int(v))
    spec = []
    for i, n in enumerate(ns):
        if tv_every and i % tv_every == 1:
            tv = _tv(n, rng)
            b.set_time_variant(i, tv)
            spec.append(("tv", tv))
        else:
            r = rng.lognormal(0.0, 0.5, n + 1)
            b.set_rates(i, r)
            spec.append(("r", r))
    return b, spec


def _oracle(n, spec, steps, h=0.01):
    x0 = oracle.uniform_x0(n)
    if spec[0] == "tv":
        ox, _ = oracle.run_rk4(n, None, x0, 0.0, h, steps, tv=spec[1])
    else:
        ox, _ = oracle.run_rk4(n, spec[1], x0, 0.0, h, steps)
    return ox


NS = [2, 5, 8, 9, 10, 12, 13, 14, 15, 16, 17]


@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_batch_matches_oracle_per_instance(prec, tol, schedule):
    """Every instance against its own oracle run, under the autotuned (0),
    one-lane-per-state (1), tile (2) and big-tile (3) schedules of the shared
    passes."""
    steps = 40
    b, spec = _setup(NS + NS[::-1], 512, prec, schedule=schedule)
    with b:
        infos = b.run(0.0, 0.01, steps)
        xs = [b.get_x(i) for i in range(b.count)]
        chosen = b.get_option(T.BATCH_OPT_SCHEDULE)
    assert chosen in (1, 2, 3) and (schedule == 0 or chosen == schedule)
    for i, n in enumerate(b.ns):
        ox = _oracle(n, spec[i], steps)
        err = np.max(np.abs(xs[i] - ox))
        print(f"schedule {chosen} instance {i} n={n} {spec[i][0]} {prec}: max|x - x_oracle| = {err:.3e}")
        assert err <= tol, (i, n, spec[i][0], err)
        assert infos[i].steps_done == steps and not infos[i].stopped
        assert infos[i].t == pytest.approx(0.01 * steps, abs=1e-12)
        assert abs(infos[i].mass - 1.0) <= (1e-12 if prec == "f64" else 1e-5)


@pytest.mark.parametrize("schedule", [1, 2, 3])
@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_batch_large_instances_match_oracle(prec, tol, schedule):
    """Instances larger than the benchmark batch's (n = 18, 20, 22: up to
    12 tile-index bits in the tile schedule) beside small ones, each against
    its own oracle run."""
    steps = 12
    b, spec = _setup([20, 9, 22, 18], 77, prec, schedule=schedule)
    with b:
        infos = b.run(0.0, 0.01, steps)
        xs = [b.get_x(i) for i in range(b.count)]
        assert b.get_option(T.BATCH_OPT_SCHEDULE) == schedule
    for i, n in enumerate(b.ns):
        ox = _oracle(n, spec[i], steps)
        err = np.max(np.abs(xs[i] - ox))
        assert err <= tol, (i, n, spec[i][0], err)
        assert infos[i].steps_done == steps


def test_batch_lds_instances_equal_single_engine_bitwise():
    """LDS-resident batch instances run the same kernel as a single context
    (RK4 keeps four vectors in LDS: n <= 12 in FP64, n <= 13 in FP32)."""
    for prec, ns in (("f64", [8, 11, 12]), ("f32", [8, 11, 13])):
        b, spec = _setup(ns, 7, prec)
        with b:
            b.run(0.0, 0.01, 30)
            xs = [b.get_x(i) for i in range(len(ns))]
        for i, n in enumerate(ns):
            with T.Engine(n, prec, "rk4", "mf") as e:
                e.set_option(T.OPT_LDS, 2)
                if spec[i][0] == "tv":
                    e.set_time_variant(spec[i][1])
                else:
                    e.set_rates(spec[i][1])
                e.set_x0_uniform()
                e.run(0.0, 0.01, 30, patience=-1)
                assert np.array_equal(e.get_x(), xs[i]), (prec, n)


def test_batch_graph_and_streams_do_not_change_results():
    """Graph replay and the second stream change nothing, bit for bit, under
    each schedule.  The schedule is fixed per group: the autotuner's choice
    between two schedules of similar speed can differ from run to run, and
    the two schedules sum rows in different orders."""
    ns = [9, 13, 14, 16, 15, 14]
    steps = 21   # two graph replays of 8 steps and 5 direct steps
    runs = []
    for opts in (dict(graph_steps=0, concurrent=0, schedule=1), dict(graph_steps=8, concurrent=1, schedule=1),
                 dict(graph_steps=8, concurrent=0, schedule=1), dict(graph_steps=8, schedule=2),
                 dict(graph_steps=0, schedule=2), dict(graph_steps=8, schedule=3),
                 dict(graph_steps=0, schedule=3), dict(graph_steps=8, schedule=3, range_tiles=1, store=1)):
        b, _ = _setup(ns, 3, "f32", **opts)
        with b:
            b.run(0.0, 0.01, steps)
            runs.append([b.get_x(i) for i in range(len(ns))])
    for r in runs[1:3]:                    # lanes: graph replay, streams
        for a, c in zip(runs[0], r):
            assert np.array_equal(a, c)
    for a, c in zip(runs[3], runs[4]):     # tiles: graph replay
        assert np.array_equal(a, c)
    for r in runs[6:]:                     # big tiles: graph replay, tile ranges, store policy
        for a, c in zip(runs[5], r):
            assert np.array_equal(a, c)


def test_batch_streamed_small_instances_match_lds():
    """With the LDS class off, n = 10..13 take the shared passes instead."""
    ns = [10, 11, 12, 13]
    out = {}
    for lds in (1, 0):
        b, spec = _setup(ns, 11, "f64", lds=lds)
        with b:
            b.run(0.0, 0.01, 30)
            out[lds] = [b.get_x(i) for i in range(len(ns))]
    for a, c in zip(out[0], out[1]):
        assert np.max(np.abs(a - c)) <= 1e-14


def test_batch_early_exit_is_per_instance():
    """alpha = beta = 1/2, gamma = 1 from the uniform state is stationary:
    those instances leave at step index 100; the others run on."""
    ns = [8, 14, 8, 14]
    b = T.Batch(ns, "f64")
    rng = np.random.default_rng(5)
    with b:
        for i, n in enumerate(ns):
            if i < 2:
                b.set_rates(i, np.array([0.5] + [1.0] * (n - 1) + [0.5]))
            else:
                b.set_rates(i, rng.uniform(0.3, 1.5, n + 1))
        infos = b.run(0.0, 0.01, 300, thr=1e-7, patience=100)
        x0 = b.get_x(0)
        x1 = b.get_x(1)
    for i in (0, 1):
        assert infos[i].stopped and infos[i].steps_at_stop == 100 and infos[i].steps_done == 101
    assert np.all(x0 == 2.0 ** -8) and np.all(x1 == 2.0 ** -14)
    for i in (2, 3):
        assert not infos[i].stopped and infos[i].steps_done == 300


def test_batch_moments_vs_oracle():
    ns = [9, 15]
    b, spec = _setup(ns, 21, "f64", tv_every=0)
    with b:
        b.run(0.0, 0.01, 25)
        for i, n in enumerate(ns):
            x = b.get_x(i)
            y, cov, _, mass = b.moments(i)
            oy = oracle.site_moments(n, x)
            ocov, _ = oracle.cov(n, x, oy)
            assert np.max(np.abs(y - oy)) <= 1e-13
            assert np.max(np.abs(cov[::-1, ::-1] - ocov)) <= 1e-13
            assert abs(mass - oracle.mass(x)) <= 1e-13
            assert np.max(np.abs(b.window_marginals(i) - oracle.windows(n, x))) <= 1e-13


def test_batch_errors():
    with pytest.raises(T.TasepError):
        T.Batch([8, 31], "f32")
    with pytest.raises(T.TasepError):
        T.Batch([], "f32")
    with T.Batch([8, 14], "f32") as b:
        b.set_rates(0, np.ones(9))
        with pytest.raises(T.TasepError):
            b.run(0.0, 0.01, 5)          # instance 1 has no rates
        with pytest.raises(T.TasepError):
            b.set_option(99, 1)
        with pytest.raises(T.TasepError):
            b.set_rates(2, np.ones(9))   # no such instance


def test_config4_workload_matches_oracle_per_instance():
    """The benchmark batch itself -- 512 instances, n ~ U{8..18}, rates
    ~ LogNormal(0, 0.5), odd instances time-variant (seed 512) -- under the
    autotuned schedule with its XCD block table and graph replay: every
    instance against its own FP64 oracle RK4 run after 20 steps (FP32
    state, <= 1e-6), every instance's steps, time and FP64 mass."""
    import bench
    ns, spec = bench.config_4_instances()
    assert len(ns) == 512 and min(ns) == 8 and max(ns) == 18
    steps, h = 20, 0.01
    with T.Batch(ns, "f32") as b:
        for i, (kind, val) in enumerate(spec):
            if kind == "tv":
                b.set_time_variant(i, val)
            else:
                b.set_rates(i, val)
        infos = b.run(0.0, h, steps)
        chosen = b.get_option(T.BATCH_OPT_SCHEDULE)
        xs = [b.get_x(i) for i in range(len(ns))]
    print(f"config-4 batch: schedule {chosen}")
    assert chosen in (1, 2, 3)
    worst = 0.0
    for i, (n, (kind, val)) in enumerate(zip(ns, spec)):
        x0 = oracle.uniform_x0(n)
        if kind == "tv":
            ox, _ = oracle.run_rk4(n, None, x0, 0.0, h, steps, tv=val)
        else:
            ox, _ = oracle.run_rk4(n, val, x0, 0.0, h, steps)
        err = float(np.max(np.abs(xs[i] - ox)))
        worst = max(worst, err)
        assert err <= 1e-6, (i, n, kind, err)
        assert infos[i].steps_done == steps and not infos[i].stopped
        assert infos[i].t == pytest.approx(h * steps, abs=1e-12)
        assert abs(infos[i].mass - 1.0) <= 1e-5, (i, infos[i].mass)
        assert abs(xs[i].sum() - 1.0) <= 1e-5
    print(f"config-4 batch: max over instances of max|x - x_oracle| = {worst:.3e}")


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_bigtile_batch_equals_single_context_tiles_bitwise(prec):
    """The big-tile schedule runs the single-lattice tile pass per instance
    range: every instance (constant rates in Horner form, time-variant ones
    in staged form) equals a single context on forced tiles bit for bit."""
    ns = [14, 18, 15, 16, 19, 14]
    steps = 13
    b, spec = _setup(ns, 31, prec, schedule=3, range_tiles=3)
    with b:
        b.run(0.0, 0.01, steps)
        assert b.get_option(T.BATCH_OPT_SCHEDULE) == 3
        xs = [b.get_x(i) for i in range(len(ns))]
    for i, n in enumerate(ns):
        with T.Engine(n, prec, "rk4", "mf") as e:
            e.set_option(T.OPT_TILED, 2)
            e.set_option(T.OPT_LDS, 0)
            if spec[i][0] == "tv":
                e.set_time_variant(spec[i][1])
            else:
                e.set_rates(spec[i][1])
            e.set_x0_uniform()
            e.run(0.0, 0.01, steps, patience=-1)
            x = e.get_x()
        print(f"instance {i} n={n} {spec[i][0]} {prec}: max|batch - context| = {np.max(np.abs(xs[i] - x)):.3e}")
        assert np.array_equal(xs[i].view(np.uint64), x.view(np.uint64)), (i, n, spec[i][0])
