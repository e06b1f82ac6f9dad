"""Pins the CPU oracle (oracle/) -- the checker the HIP path is compared to.

Pins, in order of strength:
  * golden fixtures written by the reference itself (tests/golden/):
    CSR arrays from gencrs, probabilities printed with %.20g (exact doubles),
    densities/covariances printed with %lf, stop step and final time;
  * the reference's own functions compiled from source (oracle/_ref), where
    present;
  * mathematics: closed forms (nnz, row length), column sums of a generator,
    exact stationary states, the identity k1 = k2 = k3 = k4 of the
    reference's stages.
"""
import numpy as np
import pytest

import oracle
from tests import _golden as G


# ----------------------------------------------------------------- CSR (a1)
@pytest.mark.parametrize("fx", G.csr_fixtures(), ids=lambda f: f["name"])
def test_gencrs_bit_exact_vs_reference_fixture(fx):
    rowPtr, colInd, val = oracle.gencrs(fx["n"], fx["rates"])
    assert np.array_equal(rowPtr, fx["rowPtr"])
    assert np.array_equal(colInd, fx["colInd"])
    assert np.array_equal(val.view(np.uint64), fx["val"].view(np.uint64))


@pytest.mark.parametrize("n", [2, 3, 5, 9, 12])
def test_nnz_and_rowlen_closed_forms(n):
    r = np.linspace(0.3, 1.7, n + 1)
    rowPtr, colInd, val = oracle.gencrs(n, r)
    dim = 1 << n
    assert int(rowPtr[-1]) == oracle.nnz(n) == 2 * dim + (n - 1) * dim // 4
    s = np.arange(dim, dtype=np.uint64)
    m = (s & ~(s >> np.uint64(1))) & np.uint64((1 << (n - 1)) - 1)
    pc = np.array([bin(int(v)).count("1") for v in m])
    rl = 1 + ((s >> np.uint64(n - 1)) & np.uint64(1)).astype(int) + (1 - (s & np.uint64(1)).astype(int)) + pc
    assert np.array_equal(np.diff(rowPtr).astype(int), rl)
    for st in (0, 1, dim // 2, dim - 1, dim // 3):
        assert oracle.rowlen(n, st) == rl[st]
    # columns ascending within each row (entry < diag < exit < bulk)
    for i in range(0, dim, max(1, dim // 64)):
        cols = colInd[rowPtr[i]:rowPtr[i + 1]]
        assert np.all(np.diff(cols.astype(np.int64)) > 0)


@pytest.mark.parametrize("n", [3, 6, 8])
def test_generator_columns_sum_to_zero(n):
    r = np.random.default_rng(n).uniform(0.1, 2.0, n + 1)
    L = oracle.dense(n, r)
    assert np.max(np.abs(L.sum(axis=0))) < 1e-14
    assert np.all(np.diag(L) <= 0)
    off = L - np.diag(np.diag(L))
    assert np.all(off >= 0)


def test_mf_row_equals_csr_row():
    n = 7
    r = np.random.default_rng(1).uniform(0.1, 2.0, n + 1)
    rowPtr, colInd, val = oracle.gencrs(n, r)
    for s in range(1 << n):
        c, v = oracle.row(n, s, r)
        assert np.array_equal(c, colInd[rowPtr[s]:rowPtr[s + 1]])
        assert np.array_equal(v, val[rowPtr[s]:rowPtr[s + 1]])


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built here")
def test_gencrs_and_rk4_step_vs_reference_library():
    import ctypes as C
    R = oracle.ref_lib()
    n = 9
    dim = 1 << n
    r = np.random.default_rng(3).uniform(0.1, 2.0, n + 1)
    rp = np.zeros(dim + 1, np.uint64)
    ci = np.zeros(oracle.nnz(n), np.uint64)
    va = np.zeros(oracle.nnz(n), np.float64)
    import os
    import sys
    sys.stdout.flush()
    saved = os.dup(1)
    dn = os.open(os.devnull, os.O_WRONLY)
    os.dup2(dn, 1)
    try:
        R.gencrs(rp, va, ci, r, np.zeros(n, np.int32), dim, n)
        C.CDLL(None).fflush(None)
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(dn)
    rowPtr, colInd, val = oracle.gencrs(n, r)
    assert np.array_equal(rp, rowPtr) and np.array_equal(ci, colInd)
    assert np.array_equal(va.view(np.uint64), val.view(np.uint64))
    x = np.random.default_rng(4).random(dim)
    x /= x.sum()
    k_ref = np.zeros(dim)
    x1 = np.zeros(dim)
    R.rk4_step(1, dim, k_ref, x, x1, rp, ci, va, 0.01)
    k = oracle.apply_csr(rowPtr, colInd, val, x, 0.01)
    assert np.array_equal(k.view(np.uint64), k_ref.view(np.uint64))
    assert np.array_equal(oracle.apply_mf(n, r, x, 0.01).view(np.uint64),
                          k.view(np.uint64))


# ------------------------------------------------------- step loop (a3-a6)
CLI = G.cli_case_names()


@pytest.mark.parametrize("name", CLI)
def test_run_ref_bit_exact_vs_reference_cli(name):
    c = G.cli_case(name)
    n = c["n"]
    h = oracle.step_h(c["T"], c["t0"], c["steps"])
    x0 = oracle.uniform_x0(n)
    if c["tv"] is not None:
        x, info = oracle.run_ref_tv(n, x0, c["t0"], h, c["steps"], c["tv"])
    else:
        x, info = oracle.run_ref(n, c["rates"], x0, c["t0"], h, c["steps"],
                                 use_csr=c["csr"])
    assert np.array_equal(x.view(np.uint64), c["x"].view(np.uint64)), \
        f"max|dx|={np.max(np.abs(x - c['x']))}"
    conv = G.stdout_value(c, "Convergence after ")
    if conv is not None:
        assert info.stopped or info.early_stop > 0
        assert int(conv.split()[0]) == info.steps_at_stop
    else:
        assert info.early_stop == 0
    assert G.stdout_value(c, "Simulation time elapsed:") == "%f" % info.t
    assert G.stdout_value(c, "Checksum over result vector:") == "%f" % oracle.mass(x)


@pytest.mark.parametrize("name", CLI)
def test_moments_format_vs_reference_cli(name):
    c = G.cli_case(name)
    n = c["n"]
    y = oracle.site_moments(n, c["x"])
    vec = ", ".join(["%f" % y[n - 1 - i] for i in range(n)])
    assert G.read_text(c, "out_exp_values_fin.vec") == vec
    dat = "".join("%d Pos%d %f\n" % (i, i, y[n - 1 - i]) for i in range(n))
    assert G.read_text(c, "out_exp_values_fin.dat") == dat
    cov, mul = oracle.cov(n, c["x"], y)
    mat = "".join(", ".join("%f" % v for v in row) + "\n" for row in cov)
    assert G.read_text(c, "out_cov.mat") == mat


def test_ref_stages_are_equal():
    """All four reference stages read the unmodified x: with the reference's
    own rk4_step, k1..k4 are bitwise equal, so its "RK4" update is an Euler
    step with the ((k + 2k) + 2k + k) / 6 rounding."""
    if not oracle.have_ref():
        pytest.skip("oracle/_ref not built here")
    R = oracle.ref_lib()
    n = 6
    dim = 1 << n
    r = np.random.default_rng(9).uniform(0.1, 2.0, n + 1)
    rp, ci, va = oracle.gencrs(n, r)
    x = np.random.default_rng(2).random(dim)
    x1 = np.zeros(dim)
    ks = []
    for st in (1, 2, 3, 4):
        k = np.zeros(dim)
        R.rk4_step(st, dim, k, x, x1, rp, ci, va, 0.05)
        ks.append(k)
    for k in ks[1:]:
        assert np.array_equal(k, ks[0])


# ------------------------------------------------------------ moments (a7-a9)
def test_site_moments_and_cov_identities():
    n = 7
    x = np.random.default_rng(11).random(1 << n)
    x /= x.sum()
    y = oracle.site_moments(n, x)
    s = np.arange(1 << n)
    bits = ((s[:, None] >> np.arange(n)[None, :]) & 1).astype(float)
    assert np.allclose(y, bits.T @ x, rtol=0, atol=1e-15)
    cov, mul = oracle.cov(n, x, y)
    E = (bits * x[:, None]).T @ bits
    M = x.sum()
    ident = E - np.outer(y, y) * (2 - M)      # exact identity for the centred sum
    # written orientation: cov[n-1-i][n-1-k]
    assert np.allclose(cov[::-1, ::-1], ident, rtol=0, atol=1e-14)
    assert np.allclose(mul, cov + np.outer(y, y)[::-1, ::-1], rtol=0, atol=1e-15)


# ------------------------------------------------------- rk4 + steady state
def test_rk4_oracle_matches_expm_at_fourth_order():
    from scipy.linalg import expm
    n = 5
    r = np.random.default_rng(5).uniform(0.2, 1.5, n + 1)
    L = oracle.dense(n, r)
    x0 = np.zeros(1 << n)
    x0[0] = 1.0
    exact = expm(1.0 * L) @ x0
    errs = []
    for steps in (10, 20, 40):
        x, _ = oracle.run_rk4(n, r, x0, 0.0, 1.0 / steps, steps)
        errs.append(np.max(np.abs(x - exact)))
    # fourth order: halving h divides the error by ~16
    assert errs[0] / errs[1] > 12 and errs[1] / errs[2] > 12
    assert errs[2] < 1e-6


def test_ref_loop_is_explicit_euler():
    """With k1 = k2 = k3 = k4 the reference update equals x + h L x (up to
    its 1-2-2-1 rounding)."""
    n = 5
    r = np.random.default_rng(6).uniform(0.2, 1.5, n + 1)
    L = oracle.dense(n, r)
    x0 = oracle.uniform_x0(n)
    x, _ = oracle.run_ref(n, r, x0, 0.0, 0.01, 1, patience=-1)
    assert np.allclose(x, x0 + 0.01 * (L @ x0), rtol=0, atol=1e-17)


def test_uniform_is_stationary_for_half_half_unit():
    """alpha = beta = 1/2, gamma = 1: the uniform state is exactly stationary
    (the reference stops after 101 sub-threshold steps)."""
    n = 6
    r = np.array([0.5] + [1.0] * (n - 1) + [0.5])
    x, info = oracle.run_ref(n, r, oracle.uniform_x0(n), 0.0, 0.01, 10000)
    assert info.stopped and info.steps_at_stop == 100 and info.steps_done == 101
    assert np.all(x == 1.0 / 64)


def test_dehp_matches_exact_bruteforce():
    from fractions import Fraction as F
    from oracle import dehp
    for n in (1, 2, 3, 4):
        for a, b in ((0.3, 0.7), (F(1, 3), F(3, 4)), (1.5, 0.25)):
            p = dehp.steady_state_bruteforce(n, [a] + [1] * (n - 1) + [b])
            assert dehp.densities(n, a, b) == dehp.densities_from_p(n, p)


def test_converged_ref_run_matches_dehp():
    from oracle import dehp
    n = 8
    a, b = 0.3, 0.6
    r = np.array([a] + [1.0] * (n - 1) + [b])
    x, info = oracle.run_ref(n, r, oracle.uniform_x0(n), 0, 0.02, 40000, thr=1e-10)
    assert info.stopped
    y = oracle.site_moments(n, x)[::-1]          # entry .. exit
    rho = np.array([float(v) for v in dehp.densities(n, a, b)])
    assert np.max(np.abs(y - rho)) < 1e-6


def test_gamma_rescaled_twin():
    from oracle import dehp
    n = 4
    r2 = [0.6] + [2.0] * (n - 1) + [0.8]
    p = dehp.steady_state_bruteforce(n, r2)
    assert dehp.densities_from_p(n, p) == dehp.densities(n, 0.6, 0.8, 2.0)


# ------------------------------------------------- high-order post-processing
def test_windows_equal_reference_key_sums():
    """Every key_fit sum the reference's high-order functions take (<= 4
    sites) is a sum of window marginals (oracle.key_sum restates key_fit)."""
    from tests import _highorder as H
    for n in (4, 5, 8):
        x = np.random.default_rng(60 + n).random(1 << n)
        W = oracle.windows(n, x)
        for size in (1, 2, 3, 4):
            for off in range(0, n - size + 1):
                for key in range(1 << size):
                    want = oracle.key_sum(n, x, key, off, size)
                    assert abs(H.key_sum(W, n, key, off, size) - want) <= 1e-12 * x.sum()


def test_six_site_windows_equal_key_sums():
    """6-site window marginals (the timed correlation's sums span 6 sites):
    every bin is the key_fit sum of its pattern, every key_fit sum of up to
    6 sites is a sum of bins, and dropping two edge sites gives the 4-site
    windows."""
    from tests import _highorder as H
    for n in (6, 7, 9):
        x = np.random.default_rng(70 + n).random(1 << n)
        W6 = oracle.windows(n, x, 6)
        assert W6.shape == (n - 5, 64)
        for o in range(n - 5):
            for p in range(64):
                assert abs(W6[o, p] - oracle.key_sum(n, x, p, o, 6)) <= 1e-13 * x.sum()
        for size in (4, 5, 6):
            for off in range(0, n - size + 1):
                for key in range(1 << size):
                    want = oracle.key_sum(n, x, key, off, size)
                    assert abs(H.key_sum(W6, n, key, off, size, width=6) - want) <= 1e-12 * x.sum()
        W4 = oracle.windows(n, x)
        for o in range(n - 5):
            assert np.allclose(W6[o].reshape(4, 16).sum(axis=0), W4[o], rtol=1e-13, atol=0)
            assert np.allclose(W6[o].reshape(16, 4).sum(axis=1), W4[o + 2], rtol=1e-13, atol=0)


def test_windows_invariants():
    n = 9
    x = np.random.default_rng(3).random(1 << n)
    W = oracle.windows(n, x)
    assert np.allclose(W.sum(axis=1), x.sum(), rtol=1e-14, atol=0)
    # the 3-site window o+1..o+3 from W[o] (drop bit 0) and W[o+1] (drop bit 3)
    for o in range(n - 4):
        a = W[o].reshape(8, 2).sum(axis=1)
        b = W[o + 1].reshape(2, 8).sum(axis=0)
        assert np.allclose(a, b, rtol=1e-13, atol=0)
    # neighbour pair moments E[b_o b_{o+1}] by brute force
    s = np.arange(1 << n)
    for o in range(n - 3):
        e = x[((s >> o) & 3) == 3].sum()
        assert abs(W[o][[p for p in range(16) if p & 3 == 3]].sum() - e) <= 1e-13 * x.sum()


@pytest.mark.parametrize("name", G.cli_case_names())
def test_high_order_files_from_windows_match_reference(name):
    """Error guess, closures and point functions of the reference's own runs,
    recomputed from the window marginals of its final state, text-equal."""
    from tests import _highorder as H
    c = G.cli_case(name)
    if c["n"] < 4:
        pytest.skip("no windows below n = 4")
    W = oracle.windows(c["n"], c["x"])
    for fname, text in H.files(W, c["n"]).items():
        assert text == G.read_text(c, fname), fname


# ------------------------------------------- time-variant rates and RK4 (a6)
def _tv_cases():
    """Rate functions of all three types, with nonzero offsets, amplitudes
    of both signs and constants small enough that the clamp at 0 acts."""
    rng = np.random.default_rng(44)
    m = 24
    return dict(type=np.array([0, 1, 2] * (m // 3), np.int32),
                amp=rng.uniform(-1.5, 1.5, m), period=rng.uniform(0.3, 3.0, m),
                offset=rng.uniform(-2.0, 2.0, m), c=rng.uniform(-0.5, 1.2, m))


@pytest.mark.skipif(not oracle.have_ref(), reason="oracle/_ref not built here")
def test_tv_rates_bit_exact_vs_reference_timeVariantProbs():
    """oracle.tv_rates against the reference's own timeVariantProbs
    (src/lib/matvec.c:44-68, compiled from source): sin, cos and linear
    types, the clamp at 0, nonzero offsets, and an unknown type that leaves
    the rate as it was."""
    R = oracle.ref_lib()
    tv = _tv_cases()
    m = len(tv["type"])
    clamped = 0
    for t in (0.0, 0.37, 1.0, 2.5, 7.25, -1.5, 123.456):
        want = np.zeros(m)
        R.timeVariantProbs(want, m, np.ascontiguousarray(tv["period"]), np.ascontiguousarray(tv["offset"]),
                           np.ascontiguousarray(tv["amp"]), t, tv["type"], np.ascontiguousarray(tv["c"]))
        got = oracle.tv_rates(m - 1, t, tv)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), t
        clamped += int(np.sum(want == 0.0))
        assert np.all(want >= 0)
    assert clamped > 10                               # the clamp was exercised
    # closed forms of the three types (clamp included)
    t = 0.9
    got = oracle.tv_rates(m - 1, t, tv)
    arg = t * tv["period"] + tv["offset"]
    raw = np.where(tv["type"] == 0, tv["amp"] * np.sin(arg) + tv["c"],
                   np.where(tv["type"] == 1, tv["amp"] * np.cos(arg) + tv["c"], tv["amp"] * t + tv["c"]))
    assert np.allclose(got, np.maximum(raw, 0.0), rtol=0, atol=1e-15)
    # an unknown type leaves the caller's rate untouched, as the reference's switch does
    tv3 = dict(tv, type=np.full(m, 3, np.int32))
    keep = np.full(m, 0.625)
    R.timeVariantProbs(keep, m, np.ascontiguousarray(tv["period"]), np.ascontiguousarray(tv["offset"]),
                       np.ascontiguousarray(tv["amp"]), 1.0, tv3["type"], np.ascontiguousarray(tv["c"]))
    assert np.all(keep == 0.625)


def _tv_problem(n, seed):
    rng = np.random.default_rng(seed)
    return dict(type=(np.arange(n + 1) % 3).astype(np.int32), amp=rng.uniform(0.4, 0.9, n + 1),
                period=rng.uniform(1.5, 3.0, n + 1), offset=rng.uniform(-1.0, 1.0, n + 1),
                c=rng.uniform(1.0, 1.5, n + 1))


def _exact_tv_solution(n, tv, x0, T):
    """x(T) of dx/dt = L(r(t)) x from scipy's 8th-order integrator at a tight
    tolerance; L from the oracle's dense generator (pinned above against the
    reference's gencrs), r(t) from oracle.tv_rates (pinned above against the
    reference's timeVariantProbs)."""
    from scipy.integrate import solve_ivp
    sol = solve_ivp(lambda t, x: oracle.dense(n, oracle.tv_rates(n, t, tv)) @ x, (0.0, T), x0,
                    method="DOP853", rtol=1e-13, atol=1e-16)
    assert sol.success
    return sol.y[:, -1]


def _rk4_dense(n, tv, x0, T, steps, stage_t):
    """Classical RK4 on the dense generator with stage times t + stage_t[i]*h
    (correct: 0, 1/2, 1/2, 1).  Only used to show that the convergence check
    below tells a wrong stage time apart."""
    h = T / steps
    x = x0.copy()
    t = 0.0
    for _ in range(steps):
        Ls = [oracle.dense(n, oracle.tv_rates(n, t + c * h, tv)) for c in stage_t]
        k1 = h * (Ls[0] @ x)
        k2 = h * (Ls[1] @ (x + 0.5 * k1))
        k3 = h * (Ls[2] @ (x + 0.5 * k2))
        k4 = h * (Ls[3] @ (x + k3))
        x = x + (k1 + 2 * k2 + 2 * k3 + k4) / 6
        t += h
    return x


def _order_ratios(errs):
    return [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]


def test_rk4_time_variant_oracle_is_fourth_order():
    """The oracle's classical RK4 with time-variant rates (stages at t,
    t+h/2, t+h/2, t+h) converges at fourth order to the exact solution: the
    error falls ~16x per halving of h.  The same check applied to RK4 with
    stages 2 and 3 at t or at t + h (or the first or last stage at t + h/2)
    falls only ~2x, so a stage-time mistake in the oracle fails it."""
    n = 4
    tv = _tv_problem(n, 7)
    x0 = np.zeros(1 << n)
    x0[0] = 1.0
    T = 2.0
    exact = _exact_tv_solution(n, tv, x0, T)
    steps = (20, 40, 80)
    errs = []
    for s in steps:
        x, info = oracle.run_rk4(n, None, x0, 0.0, T / s, s, tv=tv)
        assert info.steps_done == s and abs(info.t - T) < 1e-12
        errs.append(float(np.max(np.abs(x - exact))))
    ratios = _order_ratios(errs)
    assert all(q > 13 for q in ratios), (errs, ratios)
    assert errs[-1] < 5e-8
    # the dense restatement with the right stage times agrees with the oracle
    x_d = _rk4_dense(n, tv, x0, T, 40, (0.0, 0.5, 0.5, 1.0))
    x_o, _ = oracle.run_rk4(n, None, x0, 0.0, T / 40, 40, tv=tv)
    assert np.max(np.abs(x_d - x_o)) < 1e-14
    # planted mistakes (in this CPU restatement only): stages 2, 3 at t or at t + h
    # and a last or first stage at t + h/2: each of them is first order
    for wrong in ((0.0, 0.0, 0.0, 1.0), (0.0, 1.0, 1.0, 1.0), (0.0, 0.5, 0.5, 0.5),
                  (0.5, 0.5, 0.5, 1.0)):
        e = [float(np.max(np.abs(_rk4_dense(n, tv, x0, T, s, wrong) - exact))) for s in steps]
        assert max(_order_ratios(e)) < 4, (wrong, e)


# ------------------------------------------------ DEHP two-point function (f4)
def _pairs_from_p(n, p):
    """<tau_i tau_j> (sites 1-based from the entry) of a distribution indexed
    by bitmask; site i is bit n - i."""
    return {(i, j): sum(p[s] for s in range(1 << n) if (s >> (n - i)) & 1 and (s >> (n - j)) & 1)
            for i in range(1, n + 1) for j in range(i + 1, n + 1)}


def test_two_point_matches_exact_bruteforce():
    """oracle.dehp.two_point (steadystates_gmp.c:188-326 with Z_m in place of
    calculate_Z_i's Z_{m-1}, plus the exit-site case j = N) equals the pair
    moments of the exact null vector of the generator."""
    from fractions import Fraction as F
    from oracle import dehp
    for n in (2, 3, 4, 5, 6):
        for a, b in ((F(1, 3), F(3, 4)), (F(3, 2), F(1, 4)), (F(6, 10), F(9, 10))):
            p = dehp.steady_state_bruteforce(n, [a] + [1] * (n - 1) + [b])
            for (i, j), want in _pairs_from_p(n, p).items():
                assert dehp.two_point(i, j, a, b, n) == want, (n, a, b, i, j)
    # gamma-rescaled twin
    n, a, b, g = 4, F(6, 10), F(8, 10), F(2)
    p = dehp.steady_state_bruteforce(n, [a] + [g] * (n - 1) + [b])
    for (i, j), want in _pairs_from_p(n, p).items():
        assert dehp.two_point(i, j, a, b, n, g) == want


def test_two_point_current_identity():
    """Bulk current J = <tau_j (1 - tau_{j+1})> = rho_j - <tau_j tau_{j+1}> is
    the same on every bond and equals alpha (1 - rho_1) = beta rho_N (a check
    at sizes the brute force cannot reach)."""
    from fractions import Fraction as F
    from oracle import dehp
    for n, a, b in ((10, F(1, 3), F(3, 4)), (12, F(7, 5), F(2, 5)), (9, F(1, 2), F(1, 2))):
        rho = dehp.densities(n, a, b)
        J = dehp.current(n, a, b)
        assert J == b * rho[-1]
        for j in range(1, n):
            assert rho[j - 1] - dehp.two_point(j, j + 1, a, b, n) == J


def test_two_point_reference_literal_restatement():
    """The reference's own two-point sums (Z through calculate_Z_i, one index
    low): with alpha + beta = 1 the stationary state is the Bernoulli(alpha)
    product and the Z are geometric, so the literal value is alpha^2 like the
    exact one; elsewhere it is not the pair moment."""
    from fractions import Fraction as F
    from oracle import dehp
    a, b = F(3, 10), F(7, 10)
    for n in (4, 6, 9):
        for i in range(1, n - 1):
            for j in range(i + 1, n):
                assert dehp.two_point_reference(i, j, a, b, n) == a * a
                assert dehp.two_point(i, j, a, b, n) == a * a
    a, b = F(1, 3), F(3, 4)
    assert dehp.two_point_reference(2, 4, a, b, 6) != dehp.two_point(2, 4, a, b, 6)
    with pytest.raises(ValueError):
        dehp.two_point_reference(1, 6, a, b, 6)


# ------------------------------------------------------ -a snapshots (f2)
def _oracle_snapshots(c):
    """States at the reference's -a snapshots (src/masterEquation.c:646-666,
    :899-920, :950-966): after the update of every step index i with
    i % res == 0, i.e. after steps 1, res+1, 2 res+1, ... of the run."""
    n = c["n"]
    res = int(c["argv"][c["argv"].index("-a") + 1])
    h = oracle.step_h(c["T"], c["t0"], c["steps"])
    x = oracle.uniform_x0(n)
    t = c["t0"]
    snaps = []
    done = 0
    while done < c["steps"]:
        k = 1 if done == 0 else min(res, c["steps"] - done)
        if c["tv"] is not None:
            x, info = oracle.run_ref_tv(n, x, t, h, k, c["tv"])
        else:
            x, info = oracle.run_ref(n, c["rates"], x, t, h, k, patience=-1, use_csr=c["csr"])
        t = info.t
        if (done + k - 1) % res == 0:          # index of the last step run
            snaps.append(x.copy())
        done += k
    return snaps, h, res


def _matrix_blocks(snaps, n, which):
    out = []
    for j, x in enumerate(snaps):
        cov, mul = oracle.cov(n, x)
        m = cov if which == "cov" else mul
        rows = [", ".join("%f" % v for v in row) for row in m]
        out.append("%d Matrix:\n" % j + "\n".join(rows) + "\n\n\n")
    return out


ANIM = [nm for nm in G.cli_case_names() if "-a" in G.cli_case(nm)["argv"]]


def _strip_undefined(block):
    """Drop the last number of a matrix block: the snapshot covariance's
    (row 0, col 0) entry, summed onto an uninitialised accumulator by the
    reference (src/lib/matvec.c:93-105) and written last (:1263-1271)."""
    body = block.rstrip("\n")
    return body[:body.rindex(", ") + 2]


@pytest.mark.parametrize("name", ANIM)
def test_snapshot_files_vs_reference_cli(name):
    """The oracle's states at the -a snapshots reproduce the reference's
    timed expected values, snapshot probabilities and the covariance and
    multiplicative expected-value matrices of every snapshot (text-equal,
    apart from the one undefined entry per snapshot matrix)."""
    c = G.cli_case(name)
    n = c["n"]
    snaps, h, res = _oracle_snapshots(c)
    assert len(snaps) == -(-c["steps"] // res)
    times = ["%f" % (h * res * j) for j in range(len(snaps))]
    exp = "".join(times[j] + ", " + ", ".join("%f" % v for v in oracle.site_moments(n, x)[::-1]) + "\n"
                  for j, x in enumerate(snaps))
    assert exp == G.read_text(c, "out_exp_timed.txt")
    prb = "".join(times[j] + ", " + ", ".join("%f" % v for v in x) + "\n" for j, x in enumerate(snaps))
    assert prb == G.read_text(c, "out_prb.txt")
    for d, which in (("out_covariance_data", "cov"), ("out_multExpval_data", "mul")):
        ref = G.read_text(c, d + "/out_all.txt")
        want = _matrix_blocks(snaps, n, which)
        got_blocks = [b for b in ref.split("\n\n\n") if b]
        assert len(got_blocks) == len(want)
        for a, b in zip(got_blocks, want):
            assert _strip_undefined(a) == _strip_undefined(b)
