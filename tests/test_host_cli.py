"""The C host (tasep_amd/bin/masterEquation) against the reference binary's
own outputs (tests/golden/cli/*): same command line, same files.

CPU tests cover argument handling that ends before any GPU call; GPU tests
run each golden case and compare every output file.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests import _golden as G

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "tasep_amd", "bin", "masterEquation")

needs_bin = pytest.mark.skipif(not os.path.exists(BIN), reason="host binary not built")


def _run(args, cwd, stdin=""):
    return subprocess.run([BIN] + args, cwd=cwd, input=stdin, capture_output=True,
                          text=True, timeout=300)


@needs_bin
def test_fast_with_time_var_is_refused(tmp_path):
    p = _run(["-n", "4", "-f", "--time-var", "-t", "0", "-T", "1", "-s", "10"], tmp_path)
    assert p.returncode == 255
    assert "-f cannot be selected in combination with --time-var" in p.stderr


@needs_bin
def test_missing_alpha_is_refused(tmp_path):
    p = _run(["-n", "4", "--beta", "0.5", "-t", "0", "-T", "1", "-s", "10"], tmp_path)
    assert p.returncode == 255
    assert "No value for alpha found" in p.stderr


@needs_bin
def test_missing_beta_is_refused(tmp_path):
    p = _run(["-n", "4", "--alpha", "0.5", "--gamma", "1.000000000000000000000000001",
              "-t", "0", "-T", "1", "-s", "10"], tmp_path)
    # the reference aborts here with a double free; the host reports the missing beta
    assert p.returncode == 255
    assert "No value for beta found" in p.stderr


@needs_bin
def test_unknown_option_prints_usage(tmp_path):
    p = _run(["-x"], tmp_path)
    assert p.returncode == 255 and "Usage:" in p.stdout


# --------------------------------------------------------------- GPU runs
CASES = G.cli_case_names()
BYTE_EQUAL = ["out_exp_values_fin.vec", "out_exp_values_fin.dat", "out_cov.mat",
              "out_cov.ppm", "out_mulExpVa", "out_error_guess", "out_closures",
              "out_high_order_point_fun", "out_prb.txt"]


def _stdout_line(text, prefix):
    for chunk in text.replace("\r", "\n").split("\n"):
        if chunk.startswith(prefix):
            return chunk
    return None


@needs_bin
@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_host_matches_reference_files(name, tmp_path):
    c = G.cli_case(name)
    for f in os.listdir(c["dir"]):
        if f.endswith(".txt") and not f.startswith("out"):
            shutil.copy(os.path.join(c["dir"], f), tmp_path / f)
    print(f"case {name}: {' '.join(c['argv'])}", flush=True)
    p = _run(c["argv"], tmp_path)
    assert p.returncode == 0, p.stderr
    produced = sorted(f for f in os.listdir(tmp_path) if f.startswith("out"))
    expected = sorted(f for f in os.listdir(c["dir"]) if f.startswith("out"))
    assert produced == expected
    # probabilities: exact text for constant rates (bit-exact doubles)
    got = open(tmp_path / "out").read()
    ref = G.read_text(c, "out")
    if c["tv"] is None:
        assert got == ref
    else:
        gx = np.array([float(ln.split("\t\t")[1]) for ln in got.splitlines()])
        assert np.max(np.abs(gx - c["x"])) <= 1e-12
    for f in BYTE_EQUAL:
        assert open(tmp_path / f, "rb").read() == G.read_bytes(c, f), f
    # _highcorr: layout only (the reference prints an undefined value there)
    a = open(tmp_path / "out_highcorr").read().splitlines()
    b = G.read_text(c, "out_highcorr").splitlines()
    assert a[0] == b[0] and [ln.split("\t")[0] for ln in a] == [ln.split("\t")[0] for ln in b]
    for prefix in ("Convergence after", "Confidience", "Checksum over result vector",
                   "Generated inital vector"):
        want = G.stdout_value(c, prefix)
        line = _stdout_line(p.stdout, prefix)
        if want is None:
            assert line is None
        else:
            assert line is not None and line[len(prefix):].strip() == want
    sim = [ln for ln in p.stdout.splitlines() if ln.startswith("Simulation time elapsed: ")]
    assert sim and sim[-1][len("Simulation time elapsed:"):].strip() == \
        G.stdout_value(c, "Simulation time elapsed:")
    if "-a" in c["argv"]:
        _check_snapshot_outputs(c, tmp_path)


def _strip_undefined(block):
    """A snapshot matrix block without its last number: the (row 0, col 0)
    entry the reference sums onto an uninitialised accumulator
    (src/lib/matvec.c:93-105), written last (src/masterEquation.c:1263-1271)."""
    body = block.rstrip("\n")
    return body[:body.rindex(", ") + 2]


def _numbers(text):
    return np.array([float(v) for v in text.replace(",", " ").replace("Matrix:", " ").split()])


def _same_text(a, b, exact):
    """Text-equal; for time-variant runs (state within 1e-12 of the
    reference's: device sin against glibc's) the printed %lf values agree
    to their last digit, a zero's sign aside."""
    if exact:
        return a == b
    u, v = _numbers(a), _numbers(b)
    return u.shape == v.shape and float(np.max(np.abs(u - v), initial=0.0)) <= 1.5e-6


def _check_snapshot_outputs(c, out_dir):
    """-a outputs (src/masterEquation.c:1041-1060, 1097-1116, 1197-1279,
    1352-1414): timed expected values and snapshot probabilities
    byte-equal; the snapshot covariance and multiplicative expected-value
    matrices text-equal and their PPMs byte-equal apart from the one
    undefined entry per snapshot; the final PPMs byte-equal; the timed
    high-order correlation in layout (its values come from a function
    without a return value, src/masterEquation.c:47-75).  Time-variant runs:
    values to the printed digit, PPM bytes within 1."""
    exact = c["tv"] is None
    assert _same_text(open(out_dir / "out_exp_timed.txt").read(), G.read_text(c, "out_exp_timed.txt"), exact)
    for d in ("out_covariance_data", "out_multExpval_data"):
        got_files = sorted(os.listdir(out_dir / d))
        assert got_files == sorted(os.listdir(os.path.join(c["dir"], d))), d
        a = [b for b in open(out_dir / d / "out_all.txt").read().split("\n\n\n") if b]
        b = [b for b in G.read_text(c, d + "/out_all.txt").split("\n\n\n") if b]
        assert len(a) == len(b) and len(a) > 1
        for u, v in zip(a, b):
            assert _same_text(_strip_undefined(u), _strip_undefined(v), exact), d
        for f in got_files:
            if not f.endswith(".ppm"):
                continue
            u = open(out_dir / d / f, "rb").read()
            v = G.read_bytes(c, d + "/" + f)
            if f != "out_fin.ppm":        # the undefined entry's pixel comes last
                u, v = u[:-3], v[:-3]
            assert len(u) == len(v), (d, f)
            if exact:
                assert u == v, (d, f)
            else:
                du = np.frombuffer(u, np.uint8).astype(int) - np.frombuffer(v, np.uint8).astype(int)
                assert np.max(np.abs(du)) <= 1, (d, f)
    a = [ln.split() for ln in open(out_dir / "out_highcorr_timed").read().splitlines()]
    b = [ln.split() for ln in G.read_text(c, "out_highcorr_timed").splitlines()]
    assert [(r[0], len(r)) for r in a] == [(r[0], len(r)) for r in b]


@needs_bin
@pytest.mark.gpu
def test_host_snapshots_option(tmp_path):
    """-a: snapshot files exist with one row per snapshot (steps 0, a, 2a...)."""
    (tmp_path / "rates.txt").write_text("0.4 1 1 1 1 0.6")
    p = _run(["-n", "5", "-i", "rates.txt", "-t", "0", "-T", "1", "-s", "100", "-a", "10",
              "-o", "out"], tmp_path, stdin="Y\nY\n")
    assert p.returncode == 0, p.stderr
    rows = open(tmp_path / "out_exp_timed.txt").read().splitlines()
    assert len(rows) == 10
    assert rows[0].startswith("0.000000, ") and rows[1].startswith("0.100000, ")
    prb = open(tmp_path / "out_prb.txt").read().splitlines()
    assert len(prb) == 10 and len(prb[0].split(", ")) == 1 + 32
    assert os.path.isdir(tmp_path / "out_covariance_data")
    assert os.path.exists(tmp_path / "out_highcorr_timed")


@needs_bin
@pytest.mark.gpu
def test_host_timed_correlation_vs_oracle_states(tmp_path):
    """-a at n=8: the timed correlation file (4-site key, sums over 6 sites)
    comes from the device's 6-site window marginals of each snapshot; its
    values equal the correlation formula evaluated with the oracle's
    state-loop key_fit sums on the oracle's states after steps 1, 11, 21, ...
    (printed with 7 significant digits; the reference's own values are
    undefined, so only the formula is checked)."""
    import oracle
    n, steps, a = 8, 100, 10
    r = [0.4] + [1.0] * (n - 1) + [0.6]
    (tmp_path / "rates.txt").write_text(" ".join("%g" % v for v in r))
    p = _run(["-n", str(n), "-i", "rates.txt", "-t", "0", "-T", "1", "-s", str(steps), "-a", str(a),
              "-o", "out"], tmp_path, stdin="Y\nY\n")
    assert p.returncode == 0, p.stderr
    rows = [ln.split() for ln in open(tmp_path / "out_highcorr_timed").read().splitlines()]
    assert len(rows) == 10
    h = oracle.step_h(1.0, 0.0, steps)
    x = np.full(1 << n, 2.0 ** -n)
    done = 0
    for i, row in enumerate(rows):
        todo = 1 if i == 0 else a
        x, _ = oracle.run_ref(n, r, x, 0.0, h, todo, patience=-1)
        done += todo
        assert float(row[0]) == pytest.approx(h * a * i, abs=1e-6)
        vals = [float(v) for v in row[1:]]
        assert len(vals) == n - 5
        for k, got in zip(range(1, n - 4), vals):
            off = n - 4 - k
            lr = oracle.key_sum(n, x, 63, off - 1, 6)
            no = oracle.key_sum(n, x, 15, off, 4)
            left = oracle.key_sum(n, x, 31, off, 5)
            right = oracle.key_sum(n, x, 31, off - 1, 5)
            want = (lr * no - left * right) / np.sqrt(left * (no - left) * right * (no - right))
            assert abs(got - want) <= 1e-6 * abs(want) + 1e-12, (i, k, got, want)


@needs_bin
@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_host_rk4_environment_knobs(tmp_path, prec):
    """TASEP_INTEGRATOR=rk4 (and TASEP_PRECISION=f32) select classical RK4 in
    the host without new options: the printed densities (%lf, entry site
    first) equal the oracle's RK4 densities to the printed digits, and the
    probabilities file to 1e-12 (FP64) / 1e-9 (FP32)."""
    import oracle
    n, steps, T = 12, 200, 2.0
    r = np.random.default_rng(12).uniform(0.2, 1.5, n + 1)
    (tmp_path / "rates.txt").write_text(" ".join("%.17g" % v for v in r))
    env = dict(os.environ, TASEP_INTEGRATOR="rk4", TASEP_PRECISION=prec)
    p = subprocess.run([BIN, "-n", str(n), "-i", "rates.txt", "-t", "0", "-T", str(T), "-s", str(steps),
                        "-o", "out"], cwd=tmp_path, capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stderr
    h = oracle.step_h(T, 0.0, steps)
    ox, _ = oracle.run_rk4(n, r, oracle.uniform_x0(n), 0.0, h, steps)
    y = oracle.site_moments(n, ox)
    vec = [float(v) for v in open(tmp_path / "out_exp_values_fin.vec").read().split(",")]
    assert len(vec) == n
    assert max(abs(vec[i] - y[n - 1 - i]) for i in range(n)) <= 1.5e-6
    vals = np.array([float(ln.split("\t\t")[1]) for ln in open(tmp_path / "out")], np.float64)
    assert np.max(np.abs(vals - ox)) <= (1e-12 if prec == "f64" else 1e-9)


@needs_bin
@pytest.mark.gpu
@pytest.mark.parametrize("n,fast", [(14, False), (20, False), (20, True)])
def test_host_large_lattice_vs_oracle(tmp_path, n, fast):
    """Larger lattices than the fixtures (other kernels: one lane per state,
    tiles, CSR): the probabilities file equals the oracle's reference loop
    bit for bit, and the high-order files equal those of the oracle's
    window marginals."""
    import oracle
    from tests import _highorder as H
    rng = np.random.default_rng(n)
    r = rng.uniform(0.2, 1.5, n + 1)
    (tmp_path / "rates.txt").write_text(" ".join("%.17g" % v for v in r))
    steps, T = 30, 0.3
    args = ["-n", str(n), "-i", "rates.txt", "-t", "0", "-T", str(T), "-s", str(steps), "-o", "out"]
    if fast:
        args.append("-f")
    p = _run(args, tmp_path)
    assert p.returncode == 0, p.stderr
    h = oracle.step_h(T, 0.0, steps)
    ox, _ = oracle.run_ref(n, r, oracle.uniform_x0(n), 0.0, h, steps, use_csr=fast)
    vals = np.array([float(ln.split("\t\t")[1]) for ln in open(tmp_path / "out")], np.float64)
    assert np.array_equal(vals.view(np.uint64), ox.view(np.uint64))
    W = oracle.windows(n, ox)
    for fname, text in H.files(W, n).items():
        assert open(tmp_path / fname).read() == text, fname
