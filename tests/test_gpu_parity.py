"""Parity of the HIP path (through the C ABI) with the reference and the oracle.

Bars: bit-exact for the CSR arrays and for the FP64 reference-faithful loop
(TASEP_REF) with constant rates; <= 1e-12 where a libm function differs
(time-variant rates: device sin vs glibc sin) and for moments (parallel
FP64 sums vs the reference's sequential ones); stated tolerances for RK4.
"""
import numpy as np
import pytest

import oracle
import tasep_amd as T
from tests import _golden as G

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64),
                          np.asarray(b, np.float64).view(np.uint64))


# ------------------------------------------------------------------ CSR (a1)
@pytest.mark.parametrize("fx", G.csr_fixtures(), ids=lambda f: f["name"])
def test_device_csr_bit_exact_vs_reference_fixture(fx):
    rp, ci, va = T.gencrs(fx["n"], fx["rates"])
    assert np.array_equal(rp, fx["rowPtr"])
    assert np.array_equal(ci, fx["colInd"])
    assert _bits_equal(va, fx["val"])


@pytest.mark.parametrize("n", [2, 3, 13, 17])
def test_device_csr_bit_exact_vs_oracle(n):
    r = np.random.default_rng(100 + n).uniform(0.1, 2.0, n + 1)
    rp, ci, va = T.gencrs(n, r)
    orp, oci, ova = oracle.gencrs(n, r)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and _bits_equal(va, ova)
    assert int(rp[-1]) == T.csr_nnz(n)


# ---------------------------------------------------- reference loop (a2-a6)
@pytest.mark.parametrize("gen", ["mf", "csr"])
@pytest.mark.parametrize("name", G.cli_case_names())
def test_ref_loop_vs_reference_cli_outputs(name, gen):
    c = G.cli_case(name)
    if c["tv"] is not None and gen == "csr":
        pytest.skip("the reference refuses -f with --time-var")
    n = c["n"]
    h = oracle.step_h(c["T"], c["t0"], c["steps"])
    with T.Engine(n, "f64", "ref", gen) as e:
        if c["tv"] is not None:
            e.set_time_variant(c["tv"])
        else:
            e.set_rates(c["rates"])
        e.set_x0_uniform()
        info = e.run(c["t0"], h, c["steps"], thr=1e-7, patience=100)
        x = e.get_x()
        y, cov, mul, mass = e.moments()
    if c["tv"] is None:
        assert _bits_equal(x, c["x"]), f"max|dx| = {np.max(np.abs(x - c['x']))}"
    else:
        # device sin/cos vs glibc: not bit-exact, bounded
        assert np.max(np.abs(x - c["x"])) <= 1e-12
    conv = G.stdout_value(c, "Convergence after ")
    if conv is not None:
        assert info.stopped and int(conv.split()[0]) == info.steps_at_stop
    else:
        assert not info.stopped and info.steps_done == c["steps"]
    assert G.stdout_value(c, "Simulation time elapsed:") == "%f" % info.t
    assert G.stdout_value(c, "Checksum over result vector:") == "%f" % mass
    # moments vs the reference's written files (entry site first)
    vec = ", ".join("%f" % y[n - 1 - i] for i in range(n))
    assert G.read_text(c, "out_exp_values_fin.vec") == vec
    ocov, omul = oracle.cov(n, c["x"])
    assert np.max(np.abs(cov[::-1, ::-1] - ocov)) <= 1e-12
    assert np.max(np.abs(mul[::-1, ::-1] - omul)) <= 1e-12


@pytest.mark.parametrize("gen", ["mf", "csr"])
@pytest.mark.parametrize("n", [5, 12, 13, 14, 16, 19, 20])
def test_ref_loop_bit_exact_vs_oracle(n, gen):
    """The reference loop from a random state, 40 steps, bit for bit against
    the oracle; n = 19, 20 run the matrix-free loop on FP64 tiles."""
    r = np.random.default_rng(n).uniform(0.1, 2.0, n + 1)
    x0 = np.random.default_rng(n + 1).random(1 << n)
    x0 /= x0.sum()
    steps = 40
    ox, oinfo = oracle.run_ref(n, r, x0, 0.0, 0.01, steps, patience=-1)
    with T.Engine(n, "f64", "ref", gen) as e:
        e.set_rates(r)
        e.set_x(x0)
        info = e.run(0.0, 0.01, steps, patience=-1)
        x = e.get_x()
    assert info.steps_done == steps and oinfo.steps_done == steps
    assert _bits_equal(x, ox), f"max|dx| = {np.max(np.abs(x - ox))}"
    assert info.t == oinfo.t
    # the checksum: the oracle's sequential sum of 2^n terms |d| >= 0 is
    # within (2^n - 1) u of the exact sum, the device's compensated sum
    # within a few u
    tol = ((1 << n) + 2) * 2.0 ** -53
    assert abs(info.last_checksum - oinfo.last_checksum) <= tol * oinfo.last_checksum


def test_ref_loop_csr_rows_across_matrix_chunks():
    """n=20: groups of 64 rows hold ~430 matrix entries on average, so many
    groups span two 512-entry chunks of the stored-matrix kernel and rows
    continue across the chunk boundary.  Stored matrix and matrix-free rows
    must stay bitwise equal."""
    n = 20
    r = np.random.default_rng(20).uniform(0.1, 2.0, n + 1)
    xs = []
    for gen in ("csr", "mf"):
        with T.Engine(n, "f64", "ref", gen) as e:
            e.set_rates(r)
            e.set_x0_uniform()
            e.run(0.0, 0.01, 12, patience=-1)
            xs.append(e.get_x())
    assert _bits_equal(xs[0], xs[1])


def test_ref_loop_csr_at_its_largest_lattice_n30():
    """n=30, the largest lattice the stored matrix takes (9.9e9 nonzeros,
    167 GB of CSR arrays with 64-bit indices): two reference steps through
    the device-built matrix equal the matrix-free tiled loop bit for bit."""
    n = 30
    r = np.random.default_rng(30).uniform(0.1, 2.0, n + 1)
    xs = []
    for gen in ("csr", "mf"):
        with T.Engine(n, "f64", "ref", gen) as e:
            e.set_rates(r)
            e.set_x0_uniform()
            info = e.run(0.0, 0.01, 2, patience=-1)
            assert info.steps_done == 2
            xs.append(e.get_x())
    assert _bits_equal(xs[0], xs[1])


def test_ref_early_stop_n10_uniform_is_stationary():
    """n=10, alpha=beta=1/2, gamma=1 from the uniform state: the reference
    leaves after 101 sub-threshold steps (index 100, t = 1.01)."""
    n = 10
    r = np.array([0.5] + [1.0] * (n - 1) + [0.5])
    with T.Engine(n, "f64", "ref", "csr") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        info = e.run(0.0, oracle.step_h(100.0, 0.0, 10000), 10000)
        x = e.get_x()
    assert info.stopped and info.steps_at_stop == 100 and info.steps_done == 101
    assert "%f" % info.t == "1.010000"
    assert np.all(x == 2.0 ** -10)


def test_ref_n10_from_empty_lattice_vs_oracle():
    """n=10, alpha = beta = 1/2, gamma = 1, from the empty lattice
    (p[0] = 1), t = 0..100 in 10,000 steps."""
    n = 10
    r = np.array([0.5] + [1.0] * (n - 1) + [0.5])
    h = oracle.step_h(100.0, 0.0, 10000)
    x0 = np.zeros(1 << n)
    x0[0] = 1.0
    ox, oinfo = oracle.run_ref(n, r, x0, 0.0, h, 10000, use_csr=True)
    for gen in ("csr", "mf"):
        with T.Engine(n, "f64", "ref", gen) as e:
            e.set_rates(r)
            e.set_x0_state(0)
            info = e.run(0.0, h, 10000)
            x = e.get_x()
            y, cov, _, _ = e.moments()
        assert _bits_equal(x, ox)
        assert info.steps_done == oinfo.steps_done and info.stopped == oinfo.stopped
        assert np.max(np.abs(y - oracle.site_moments(n, ox))) <= 1e-12


# ------------------------------------------------------------ moments (a7-a9)
@pytest.mark.parametrize("n", [2, 7, 10, 11, 15, 18])
def test_moments_vs_oracle(n):
    x = np.random.default_rng(7 * n).random(1 << n)
    x /= x.sum()
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_x(x)
        y, cov, mul, mass = e.moments()
    oy = oracle.site_moments(n, x)
    ocov, omul = oracle.cov(n, x, oy)
    assert np.max(np.abs(y - oy)) <= 1e-13
    assert abs(mass - oracle.mass(x)) <= 1e-13
    assert np.max(np.abs(cov[::-1, ::-1] - ocov)) <= 1e-13
    assert np.max(np.abs(mul[::-1, ::-1] - omul)) <= 1e-13


@pytest.mark.parametrize("n", [4, 5, 9, 10, 11, 13, 14, 17, 20])
def test_window_marginals_vs_oracle(n):
    x = np.random.default_rng(90 + n).random(1 << n)
    x /= x.sum()
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_x(x)
        W = e.window_marginals()
    assert np.max(np.abs(W - oracle.windows(n, x))) <= 1e-13
    # FP32 contexts: divided by the FP64 mass of the stored vector
    with T.Engine(n, "f32", "rk4", "mf") as e:
        e.set_x(2.0 * x)
        W32 = e.window_marginals()
    assert np.max(np.abs(W32 - oracle.windows(n, x))) <= 1e-7


@pytest.mark.parametrize("n", [6, 9, 10, 11, 13, 16, 17, 20])
def test_six_site_window_marginals_vs_oracle(n):
    """6-site windows (row-internal, row-straddling and row-index offsets of
    the device pass for n up to 20) against the oracle, and their 4-site
    marginals against the device's 4-site table."""
    x = np.random.default_rng(130 + n).random(1 << n)
    x /= x.sum()
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_x(x)
        W6 = e.window_marginals(6)
        W4 = e.window_marginals()
    assert np.max(np.abs(W6 - oracle.windows(n, x, 6))) <= 1e-13
    for o in range(n - 5):
        assert np.max(np.abs(W6[o].reshape(4, 16).sum(axis=0) - W4[o])) <= 1e-13
    with T.Engine(n, "f32", "rk4", "mf") as e:
        e.set_x(2.0 * x)
        W32 = e.window_marginals(6)
    assert np.max(np.abs(W32 - oracle.windows(n, x, 6))) <= 1e-7


# ------------------------------------------------------------------- RK4 mode
@pytest.mark.parametrize("gen", ["mf", "csr"])
def test_rk4_f64_vs_oracle(gen):
    n = 11
    r = np.random.default_rng(3).uniform(0.2, 1.5, n + 1)
    x0 = np.zeros(1 << n)
    x0[0] = 1.0
    ox, _ = oracle.run_rk4(n, r, x0, 0.0, 0.01, 200)
    with T.Engine(n, "f64", "rk4", gen) as e:
        e.set_rates(r)
        e.set_x0_state(0)
        info = e.run(0.0, 0.01, 200, patience=-1)
        x = e.get_x()
    assert info.steps_done == 200
    assert np.max(np.abs(x - ox)) <= 1e-12


@pytest.mark.parametrize("n,tv", [(20, False), (20, True), (22, False)])
def test_rk4_f64_tiles_vs_oracle(n, tv):
    """RK4 on FP64 tiles (Horner passes for constant rates, stored stages for
    time-variant ones) against the oracle's classical RK4 from a random
    state, 12 steps, every state within 1e-12 of the largest probability
    (the two sum in different orders; rounding stays near 1e-15)."""
    rng = np.random.default_rng(200 + n)
    r = rng.uniform(0.2, 1.5, n + 1)
    tvd = dict(type=(np.arange(n + 1) % 3).astype(np.int32), amp=rng.uniform(0, 0.5, n + 1),
               period=rng.uniform(0.5, 2, n + 1), offset=rng.uniform(0, 1, n + 1), c=np.ones(n + 1))
    x0 = rng.random(1 << n)
    x0 /= x0.sum()
    steps = 12
    ox, _ = oracle.run_rk4(n, None if tv else r, x0, 0.0, 0.01, steps, tv=tvd if tv else None)
    with T.Engine(n, "f64", "rk4", "mf") as e:
        if tv:
            e.set_time_variant(tvd)
        else:
            e.set_rates(r)
        e.set_x(x0)
        info = e.run(0.0, 0.01, steps, patience=-1)
        x = e.get_x()
    assert info.steps_done == steps
    assert np.max(np.abs(x - ox)) <= 1e-12 * np.max(np.abs(ox))


def test_rk4_f64_time_variant_vs_oracle():
    n = 9
    rng = np.random.default_rng(4)
    tv = dict(type=np.array([0, 1, 2] * 3 + [0], np.int32)[: n + 1],
              amp=rng.uniform(0, 0.5, n + 1), period=rng.uniform(0.5, 2, n + 1),
              offset=rng.uniform(0, 1, n + 1), c=np.ones(n + 1))
    x0 = oracle.uniform_x0(n)
    ox, _ = oracle.run_rk4(n, None, x0, 0.0, 0.01, 150, tv=tv)
    with T.Engine(n, "f64", "rk4", "mf") as e:
        e.set_time_variant(tv)
        e.set_x0_uniform()
        info = e.run(0.0, 0.01, 150)
        x = e.get_x()
    assert info.steps_done == 150
    assert np.max(np.abs(x - ox)) <= 1e-12


def test_rk4_f32_densities_vs_f64_oracle():
    n = 14
    rng = np.random.default_rng(28)
    r = np.concatenate([[rng.uniform(0.2, 1.0)], rng.lognormal(0, 0.3, n - 1),
                        [rng.uniform(0.2, 1.0)]])
    x0 = oracle.uniform_x0(n)
    ox, _ = oracle.run_rk4(n, r, x0, 0.0, 0.01, 500)
    oy = oracle.site_moments(n, ox)
    ocov, _ = oracle.cov(n, ox, oy)
    with T.Engine(n, "f32", "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        e.run(0.0, 0.01, 500)
        y, cov, _, mass = e.moments()
    assert np.max(np.abs(y - oy)) <= 1e-6
    assert np.max(np.abs(cov[::-1, ::-1] - ocov)) <= 1e-6


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rk4_bernoulli_product_is_stationary_large(prec):
    """alpha + beta = 1, gamma = 1: the Bernoulli(alpha) product measure is
    the exact stationary state, at any n (a size-independent property)."""
    n = 22
    a = 0.25
    r = np.array([a] + [1.0] * (n - 1) + [1 - a])
    s = np.arange(1 << n, dtype=np.uint64)
    occ = np.zeros(1 << n)
    for k in range(n):
        occ += ((s >> np.uint64(k)) & np.uint64(1)).astype(np.float64)
    x0 = a ** occ * (1 - a) ** (n - occ)
    with T.Engine(n, prec, "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x(x0)
        info = e.run(0.0, 0.01, 20, patience=-1)
        y, cov, _, mass = e.moments()
    tol = 1e-6 if prec == "f32" else 1e-12
    assert np.max(np.abs(y - a)) <= tol
    assert np.max(np.abs(cov - np.diag(np.diag(cov)))) <= tol
    assert info.last_checksum <= (1e-4 if prec == "f32" else 1e-12)


def _bernoulli_product(n, a):
    """x_s = a^N(s) (1-a)^(n-N(s)), N(s) = occupied sites: a Kronecker
    product of n identical site marginals."""
    x = np.array([1.0])
    site = np.array([1.0 - a, a])
    for _ in range(n):
        x = np.kron(x, site)
    return x


@pytest.mark.parametrize("kind", ["horner", "staged-tv"])
def test_full_size_n28_bernoulli_product_is_stationary(kind):
    """The n=28 benchmark size (2^28 states, FP32 tiles with the per-XCD
    sweep and pre-issued streams): alpha + beta = 1, gamma = 1 keeps the
    Bernoulli(alpha) product measure stationary.  The staged passes run on
    time-variant rates that are constant in time (linear, zero slope)."""
    n, a = 28, 0.3
    r = np.array([a] + [1.0] * (n - 1) + [1 - a])
    with T.Engine(n, "f32", "rk4", "mf") as e:
        if kind == "horner":
            e.set_rates(r)
        else:
            e.set_time_variant(dict(type=np.full(n + 1, 2, np.int32), amp=np.zeros(n + 1),
                                    period=np.ones(n + 1), offset=np.zeros(n + 1), c=r))
        e.set_x(_bernoulli_product(n, a))
        info = e.run(0.0, 0.01, 6, patience=-1)
        y, cov, _, mass = e.moments()
    assert info.steps_done == 6
    assert np.max(np.abs(y - a)) <= 1e-6
    assert np.max(np.abs(cov - np.diag(np.diag(cov)))) <= 1e-6
    assert abs(mass - 1.0) <= 1e-5


@pytest.mark.parametrize("n,prec", [(32, "f32"), (28, "f64"), (34, "f32")])
def test_full_size_uniform_is_stationary(n, prec):
    """alpha = beta = 1/2, gamma = 1: the uniform distribution is exactly
    stationary (every state's out-rate equals its in-rate), at the largest
    benchmark size (n=32, 2^32 FP32 states, 16 GiB per vector), at n=28 in
    FP64 and at the largest lattice a context takes (n=34, 64 GiB per FP32
    vector, 192 GiB for the three Horner vectors)."""
    r = np.array([0.5] + [1.0] * (n - 1) + [0.5])
    with T.Engine(n, prec, "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        info = e.run(0.0, 0.01, 3, patience=-1)
        y, cov, _, mass = e.moments()
    tol = 1e-6 if prec == "f32" else 1e-12
    assert info.steps_done == 3
    assert np.max(np.abs(y - 0.5)) <= tol
    assert np.max(np.abs(cov - np.diag(np.diag(cov)))) <= tol
    assert abs(mass - 1.0) <= (1e-5 if prec == "f32" else 1e-12)


def test_ref_loop_largest_fp64_lattice_n33_uniform_is_stationary():
    """The reference loop on FP64 tiles at n=33 (64 GiB per vector, 21
    tile-index bits): alpha = beta = 1/2, gamma = 1 leaves the uniform state
    exactly unchanged -- every row sums exact multiples of x to 0 -- so the
    checksum is exactly 0 and the densities are 1/2."""
    n = 33
    r = np.array([0.5] + [1.0] * (n - 1) + [0.5])
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        info = e.run(0.0, 0.01, 3, patience=-1)
        y, cov, _, mass = e.moments()
    assert info.steps_done == 3 and info.last_checksum == 0.0
    assert np.max(np.abs(y - 0.5)) <= 1e-12
    assert abs(mass - 1.0) <= 1e-12


def _product_np(n, p):
    """numpy twin of tasep_set_x0_product: prod over bits k ascending of
    (b_k ? p[k] : 1 - p[k]), the same FP64 operations in the same order."""
    s = np.arange(1 << n, dtype=np.uint64)
    x = np.ones(1 << n)
    for k in range(n):
        b = ((s >> np.uint64(k)) & np.uint64(1)).astype(bool)
        x *= np.where(b, p[k], 1.0 - p[k])
    return x


def _site_densities(n, seed):
    """Distinct densities per bit (the state's value depends on every bit)."""
    return np.random.default_rng(seed).uniform(0.15, 0.85, n)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_product_state_and_range_read(prec):
    n = 12
    p = _site_densities(n, 3)
    want = _product_np(n, p)
    with T.Engine(n, prec, "rk4", "mf") as e:
        e.set_x0_product(p)
        x = e.get_x()
        part = e.get_x_range(1000, 77)
        assert e.get_x_range(4095, 1).shape == (1,)
        with pytest.raises(T.TasepError):
            e.get_x_range(4090, 7)
        with pytest.raises(T.TasepError):
            e.set_x0_product(np.full(n, 1.5))
    if prec == "f64":
        assert _bits_equal(x, want)
    else:
        assert np.array_equal(x, want.astype(np.float32).astype(np.float64))
    assert np.array_equal(part, x[1000:1077])


def _euler_densities(q, r, h):
    """Site densities (entry .. exit) after one explicit Euler step from the
    product measure with site densities q: rho_j' = q_j + h (J_{j-1} - J_j),
    J_0 = alpha (1 - q_1), J_j = r_j q_j (1 - q_{j+1}), J_n = beta q_n
    (the product measure factorises every pair moment)."""
    n = len(q)
    J = np.empty(n + 1)
    J[0] = r[0] * (1 - q[0])
    J[n] = r[n] * q[n - 1]
    for j in range(1, n):
        J[j] = r[j] * q[j - 1] * (1 - q[j])
    return q + h * (J[:-1] - J[1:])


def test_ref_loop_n33_product_state_one_step_closed_form():
    """The reference loop (an Euler step, four equal stages) on FP64 tiles
    at n=33 (64 GiB per vector, 2^21 tiles, the 16384-workgroup grid) from a
    product state whose value depends on every bit, with random rates: the
    densities after one step equal the closed form to rounding, and the
    mass stays 1.  A partner read from a wrong address (a truncated index, a
    wrong tile, chunk or lane) changes some density by about h times the
    affected mass."""
    n, h = 33, 0.01
    pb = _site_densities(n, 33)
    q = pb[::-1]                                    # site densities entry .. exit
    r = np.random.default_rng(34).uniform(0.2, 1.8, n + 1)
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_rates(r)
        e.set_x0_product(pb)
        y0, _, _, m0 = e.moments()
        info = e.run(0.0, h, 1, patience=-1)
        y, _, _, mass = e.moments()
    assert np.max(np.abs(y0 - pb)) <= 1e-13 and abs(m0 - 1.0) <= 1e-13
    want = _euler_densities(q, r, h)
    err = np.max(np.abs(y[::-1] - want))
    print(f"n=33 one reference step: max |rho - closed form| = {err:.3e}, |mass - 1| = {abs(mass - 1):.2e}")
    assert info.steps_done == 1
    assert err <= 1e-13
    assert abs(mass - 1.0) <= 1e-13


def _slices(dim, count=24, width=1 << 14, seed=0):
    """Offsets of state slices spread over the whole vector: both ends, the
    half and quarter points (the top bits), and random positions."""
    rng = np.random.default_rng(seed)
    offs = {0, dim - width, dim // 2 - width // 2, dim // 4, 3 * dim // 4, dim // 8 - width}
    while len(offs) < count:
        offs.add(int(rng.integers(0, dim - width)))
    return sorted(offs), width


def _rk4_product_run(n, prec, r, pb, steps, slices, width, **opts):
    with T.Engine(n, prec, "rk4", "mf") as e:
        for k, v in opts.items():
            e.set_option(getattr(T, "OPT_" + k.upper()), int(v))
        e.set_rates(r)
        e.set_x0_product(pb)
        info = e.run(0.0, 0.01, steps, patience=-1)
        y, cov, _, mass = e.moments()
        W = e.window_marginals()
        xs = [e.get_x_range(o, width) for o in slices]
    assert info.steps_done == steps
    return y, cov, mass, W, xs


def test_config5_n32_rk4_f64_tiles_vs_per_state_kernel_from_product_state():
    """Config 5 (n=32, alpha = 0.3, beta = 0.7, gamma = 1) in FP64, from a
    product state whose value depends on every bit: two RK4 steps on the
    tiles (16384 workgroups, per-XCD sweep) against the one-lane-per-state
    kernel, which derives every neighbour from the bitmask of the global
    index.  Compared: moments and 4-site window marginals over all 2^32
    states, and 24 slices of 2^14 states element by element."""
    n = 32
    r = np.array([0.3] + [1.0] * (n - 1) + [0.7])
    pb = _site_densities(n, 32)
    offs, width = _slices(1 << n)
    a = _rk4_product_run(n, "f64", r, pb, 2, offs, width)
    b = _rk4_product_run(n, "f64", r, pb, 2, offs, width, tiled=0)
    assert np.max(np.abs(a[0] - b[0])) <= 1e-13
    assert np.max(np.abs(a[1] - b[1])) <= 1e-13
    assert abs(a[2] - b[2]) <= 1e-13
    assert np.max(np.abs(a[3] - b[3])) <= 1e-13
    worst = max(float(np.max(np.abs(u - v) / np.abs(v))) for u, v in zip(a[4], b[4]))
    print(f"n=32 FP64 tiles vs one lane: worst relative element difference {worst:.2e}")
    assert worst <= 1e-12


def test_config5_n32_rk4_f32_tiles_vs_f64_tiles_from_product_state():
    """The config-5 FP32 run (n=32, 2^32 FP32 states) against the FP64 tiles
    (themselves checked against the one-lane kernel above) from the same
    product state, two steps: densities, covariances and window marginals
    within 1e-6, every element of 24 slices within 1e-5 relative."""
    n = 32
    r = np.array([0.3] + [1.0] * (n - 1) + [0.7])
    pb = _site_densities(n, 32)
    offs, width = _slices(1 << n)
    a = _rk4_product_run(n, "f32", r, pb, 2, offs, width)
    b = _rk4_product_run(n, "f64", r, pb, 2, offs, width)
    assert np.max(np.abs(a[0] - b[0])) <= 1e-6
    assert np.max(np.abs(a[1] - b[1])) <= 1e-6
    assert abs(a[2] - b[2]) <= 1e-5
    assert np.max(np.abs(a[3] - b[3])) <= 1e-6
    worst = max(float(np.max(np.abs(u - v) / np.abs(v))) for u, v in zip(a[4], b[4]))
    print(f"n=32 FP32 vs FP64 tiles: worst relative element difference {worst:.2e}")
    assert worst <= 1e-5


@pytest.mark.parametrize("n,steps", [(20, 30), (22, 12)])
def test_tiled_f32_vs_oracle_from_product_state(n, steps):
    """FP32 tiled passes at n = 20, 22 (several tile bonds per tile, the
    reads beyond the pre-issued one) against the FP64 oracle, from a product
    state whose value depends on every bit: every element within 1e-4
    relative (a wrong partner moves an element by about h r, 1e-2
    relative)."""
    r = np.random.default_rng(60 + n).uniform(0.2, 1.5, n + 1)
    pb = _site_densities(n, 61 + n)
    x0 = _product_np(n, pb)
    ox, _ = oracle.run_rk4(n, r, x0, 0.0, 0.01, steps)
    with T.Engine(n, "f32", "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_product(pb)
        e.run(0.0, 0.01, steps, patience=-1)
        x = e.get_x()
    rel = float(np.max(np.abs(x - ox) / np.abs(ox)))
    print(f"n={n} FP32 tiles vs oracle: worst relative element error {rel:.2e}")
    assert rel <= 1e-4


def _random_state(n, seed):
    x = np.random.default_rng(seed).random(1 << n)
    return x / x.sum()


def test_grid_16384_tile_orders_bitwise_n27_ref_loop():
    """The 16384-workgroup grid (two FP64 tiles per workgroup at n=27) in
    every tile order, from a random normalised state: the tiled reference
    loop equals the one-lane kernel bit for bit."""
    n = 27
    r = np.random.default_rng(27).uniform(0.2, 1.8, n + 1)
    x0 = _random_state(n, 270)
    outs = []
    for opts in [dict(tiled=0)] + [dict(tiled=1, tile_blocks=16384, xcd_order=o) for o in range(5)]:
        with T.Engine(n, "f64", "ref", "mf") as e:
            for k, v in opts.items():
                e.set_option(getattr(T, "OPT_" + k.upper()), v)
            e.set_rates(r)
            e.set_x(x0)
            e.run(0.0, 0.01, 2, patience=-1)
            outs.append(e.get_x())
    for o in outs[1:]:
        assert _bits_equal(o, outs[0])


def test_grid_16384_tile_orders_n27_rk4_f64():
    """The same grid and orders for the RK4 tiles (FP64 Horner): every
    order gives the same bits, and the tiles agree with the one-lane kernel
    to 1e-13 relative per element."""
    n = 27
    r = np.random.default_rng(28).uniform(0.2, 1.8, n + 1)
    x0 = _random_state(n, 271)
    outs = []
    for opts in [dict(tiled=0)] + [dict(tiled=1, tile_blocks=16384, xcd_order=o) for o in range(5)]:
        with T.Engine(n, "f64", "rk4", "mf") as e:
            for k, v in opts.items():
                e.set_option(getattr(T, "OPT_" + k.upper()), v)
            e.set_rates(r)
            e.set_x(x0)
            e.run(0.0, 0.01, 2, patience=-1)
            outs.append(e.get_x())
    for o in outs[2:]:
        assert _bits_equal(o, outs[1])
    assert float(np.max(np.abs(outs[1] - outs[0]) / np.abs(outs[0]))) <= 1e-13


@pytest.mark.parametrize("n,prec", [(28, "f32"), (30, "f32"), (29, "f64")])
def test_default_grid_bitwise_vs_1024_workgroups(n, prec):
    """The default grid of the RK4 tile passes -- one workgroup per CU up to
    2^15 tiles (FP32 n=28: 128 tiles per workgroup, three walk-table
    chunks), 4096 workgroups at 2^17 tiles (FP32 n=30, FP64 n=29) -- gives
    the bits of 1024 workgroups: two steps from a product state whose value
    depends on every bit; moments, window marginals and 12 state slices
    spread over the vector compared bit for bit."""
    r = np.random.default_rng(90 + n).uniform(0.2, 1.8, n + 1)
    pb = _site_densities(n, 91 + n)
    offs, width = _slices(1 << n, count=12)
    a = _rk4_product_run(n, prec, r, pb, 2, offs, width)
    b = _rk4_product_run(n, prec, r, pb, 2, offs, width, tile_blocks=1024)
    for u, v in zip(a[:2], b[:2]):
        assert _bits_equal(u, v)
    assert a[2] == b[2]
    assert _bits_equal(a[3], b[3])
    for u, v in zip(a[4], b[4]):
        assert _bits_equal(u, v)


@pytest.mark.parametrize("prec,integ", [("f64", "ref"), ("f32", "rk4")])
def test_level_walk_table_chunks_n27(prec, integ):
    """The level walk with 256 workgroups at n=27: 128 FP64 / 64 FP32 tiles
    per workgroup, so each workgroup reloads its part of the walk table once
    or twice (a chunk holds 63 tiles).  From a random normalised state the
    result equals the grid order's bit for bit: a tile taken twice or
    skipped would leave a stale or doubled slice."""
    n = 27
    r = np.random.default_rng(2727).uniform(0.2, 1.8, n + 1)
    x0 = _random_state(n, 272)
    outs = []
    for order, blocks in ((0, 256), (4, 256)):
        with T.Engine(n, prec, integ, "mf") as e:
            e.set_option(T.OPT_TILED, 1)
            e.set_option(T.OPT_TILE_BLOCKS, blocks)
            e.set_option(T.OPT_XCD_ORDER, order)
            e.set_rates(r)
            e.set_x(x0)
            e.run(0.0, 0.01, 2, patience=-1)
            outs.append(e.get_x())
    assert _bits_equal(outs[0], outs[1])
    assert not _bits_equal(outs[1], x0)


def test_full_size_n28_ref_tiles_bitwise_vs_per_state_kernel():
    """The bit-exact reference loop at the n=28 benchmark size: the tiled
    kernel (per-XCD sweep, pre-issued streams, hoisted diagonal, division by
    6 through the FMA correction) equals the one-lane-per-state kernel bit
    for bit."""
    n = 28
    r = np.random.default_rng(28).uniform(0.1, 2.0, n + 1)
    xs, infos = [], []
    for tiled in (1, 0):
        with T.Engine(n, "f64", "ref", "mf") as e:
            e.set_option(T.OPT_TILED, tiled)
            e.set_rates(r)
            e.set_x0_uniform()
            infos.append(e.run(0.0, 0.01, 3, patience=-1))
            xs.append(e.get_x())
    assert _bits_equal(xs[0], xs[1])
    assert abs(infos[0].last_checksum - infos[1].last_checksum) <= 1e-14 * infos[1].last_checksum


def test_converged_rk4_matches_dehp():
    from oracle import dehp
    n = 12
    a, b = 0.35, 0.6
    r = np.array([a] + [1.0] * (n - 1) + [b])
    with T.Engine(n, "f64", "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        info = e.run(0.0, 0.05, 40000, thr=1e-10, patience=100)
        y, _, _, _ = e.moments()
    assert info.stopped
    rho = np.array([float(v) for v in dehp.densities(n, a, b)])
    assert np.max(np.abs(y[::-1] - rho)) < 1e-7


def test_converged_pair_moments_match_dehp_two_point():
    """Converged FP64 RK4 pair moments <tau_i tau_j> (all i < j, n=12)
    against the exact two-point function (steadystates_gmp.c:188-326 with
    the Z index of the formula; pinned by brute force on the CPU)."""
    from oracle import dehp
    n = 12
    a, b = 0.35, 0.6
    r = np.array([a] + [1.0] * (n - 1) + [b])
    with T.Engine(n, "f64", "rk4", "mf") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        info = e.run(0.0, 0.05, 40000, thr=1e-11, patience=100)
        y, cov, mul, mass = e.moments()
    assert info.stopped
    assert abs(mass - 1.0) < 1e-12
    err = 0.0
    for i in range(1, n + 1):
        for j in range(i + 1, n + 1):
            want = float(dehp.two_point(i, j, a, b, n))
            err = max(err, abs(mul[n - i, n - j] - want), abs(mul[n - j, n - i] - want))
    print(f"n={n} max |<tau_i tau_j> - DEHP| = {err:.3e}")
    assert err < 1e-8


def _config2_twin():
    """Homogeneous twin of the n=20 configuration with rates ~ U[0.1, 2]
    (seed 20): its alpha and beta with every bulk rate 1 (alpha = 0.632,
    beta = 0.989, maximal-current phase)."""
    n = 20
    r = np.random.default_rng(20).uniform(0.1, 2.0, n + 1)
    return n, float(r[0]), float(r[n])


@pytest.mark.parametrize("integ,prec,h", [("ref", "f64", 0.05), ("rk4", "f64", 0.1)])
def test_config2_twin_stationary_current_matches_dehp(integ, prec, h):
    """n=20 homogeneous twin run to convergence (exit test at 1e-12 h): the
    per-site densities and the stationary current, taken at the entry, at the
    exit and on every bulk bond (rho_j - <tau_j tau_{j+1}>), equal the exact
    DEHP steady state."""
    from oracle import dehp
    n, a, b = _config2_twin()
    r = np.array([a] + [1.0] * (n - 1) + [b])
    with T.Engine(n, prec, integ, "mf") as e:
        e.set_rates(r)
        e.set_x0_uniform()
        info = e.run(0.0, h, 400000, thr=1e-12, patience=100)
        y, cov, mul, mass = e.moments()
    print(f"{integ}: stopped after {info.steps_done} steps, t = {info.t:.1f}")
    assert info.stopped
    rho = np.array([float(v) for v in dehp.densities(n, a, b)])
    site = y[::-1] / mass                                 # entry .. exit
    assert np.max(np.abs(site - rho)) < 1e-8
    J = float(dehp.current(n, a, b))
    Js = [a * (1 - site[0]), b * site[-1]]
    Js += [site[j - 1] - mul[n - j, n - j - 1] / mass for j in range(1, n)]
    print(f"DEHP J = {J:.12f}, max |J_k - J| = {max(abs(v - J) for v in Js):.3e}")
    assert max(abs(v - J) for v in Js) < 1e-8


# ----------------------------------------------------- tiled kernel coverage
def _rk4_run(n, prec, r, steps, **opts):
    opts.setdefault("lds", 0)   # these tests cover the streaming kernels
    with T.Engine(n, prec, "rk4", "mf") as e:
        for k, v in opts.items():
            e.set_option(getattr(T, "OPT_" + k.upper()), int(v))
        e.set_rates(r)
        e.set_x0_state(0)
        e.run(0.0, 0.01, steps, patience=-1)
        return e.get_x()


@pytest.mark.parametrize("n", [12, 13, 14, 15, 17])
def test_tiled_f64_matches_per_state_kernel(n):
    r = np.random.default_rng(40 + n).uniform(0.2, 1.5, n + 1)
    xt = _rk4_run(n, "f64", r, 60, tiled=2)
    xs = _rk4_run(n, "f64", r, 60, tiled=0)
    assert np.max(np.abs(xt - xs)) <= 1e-14


@pytest.mark.parametrize("n", [13, 14, 15, 16, 18])
def test_tiled_f32_matches_f64_oracle(n):
    r = np.random.default_rng(50 + n).uniform(0.2, 1.5, n + 1)
    x0 = np.zeros(1 << n)
    x0[0] = 1.0
    ox, _ = oracle.run_rk4(n, r, x0, 0.0, 0.01, 60)
    xt = _rk4_run(n, "f32", r, 60, tiled=2)
    err = np.max(np.abs(xt - ox))
    print(f"n={n} FP32 tiled max|x - x_f64| = {err:.3e}")
    assert err <= 1e-6
    assert abs(xt.sum() - 1.0) <= 1e-5


def test_xcd_order_does_not_change_results():
    n = 17
    r = np.random.default_rng(7).uniform(0.2, 1.5, n + 1)
    a = _rk4_run(n, "f32", r, 30, tiled=2, xcd_order=1)
    b = _rk4_run(n, "f32", r, 30, tiled=2, xcd_order=0)
    c = _rk4_run(n, "f32", r, 30, tiled=2, tile_blocks=64)
    assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("case", ["f32-horner", "f32-staged-tv", "f64-horner", "f64-ref"])
def test_tile_orders_bitwise_equal(case):
    """Every tile order (grid, XCD eighths, per-XCD sweep, device sweep,
    level walk), grid size and output store policy gives the same bits: each state's row
    is summed in a fixed order whichever workgroup owns its tile.  n=22 with
    256 workgroups (also the default grid there) gives every workgroup
    several tiles, so the sweeps are really taken; 1024 gives each FP64
    workgroup one tile (FP32: 512 workgroups, one tile each)."""
    n = 22
    prec, kind = case.split("-", 1)
    rng = np.random.default_rng(23)
    r = rng.uniform(0.2, 1.5, n + 1)
    tv = dict(type=(np.arange(n + 1) % 3).astype(np.int32), amp=rng.uniform(0, 0.5, n + 1),
              period=rng.uniform(0.5, 2, n + 1), offset=rng.uniform(0, 1, n + 1), c=np.ones(n + 1))
    out = []
    for order, blocks, store in ((1, 0, 0), (0, 256, 0), (1, 256, 0), (2, 256, 0), (3, 256, 0), (2, 0, 0),
                                 (3, 0, 0), (2, 0, 1), (2, 0, 2), (2, 1024, 0), (1, 1024, 0), (4, 0, 0),
                                 (4, 256, 0), (4, 1024, 0), (4, 0, 1)):
        with T.Engine(n, prec, "ref" if kind == "ref" else "rk4", "mf") as e:
            e.set_option(T.OPT_XCD_ORDER, order)
            e.set_option(T.OPT_TILE_BLOCKS, blocks)
            e.set_option(T.OPT_STORE, store)
            if kind == "staged-tv":
                e.set_time_variant(tv)
            else:
                e.set_rates(r)
            e.set_x0_uniform()
            e.run(0.0, 0.01, 12, patience=-1)
            out.append(e.get_x())
    for o in out[1:]:
        assert _bits_equal(out[0], o)


@pytest.mark.parametrize("prec,tiled", [("f64", 2), ("f64", 0), ("f32", 2)])
def test_horner_equals_staged_rk4(prec, tiled):
    n = 15
    r = np.random.default_rng(8).uniform(0.2, 1.5, n + 1)
    a = _rk4_run(n, prec, r, 50, tiled=tiled, rk4_staged=0)
    b = _rk4_run(n, prec, r, 50, tiled=tiled, rk4_staged=1)
    assert np.max(np.abs(a - b)) <= (1e-14 if prec == "f64" else 2e-7)


def test_time_variant_tiled_vs_oracle():
    n = 15
    rng = np.random.default_rng(12)
    tv = dict(type=np.zeros(n + 1, np.int32), amp=rng.uniform(0, 0.5, n + 1),
              period=rng.uniform(0.5, 2, n + 1), offset=np.zeros(n + 1), c=np.ones(n + 1))
    ox, _ = oracle.run_rk4(n, None, oracle.uniform_x0(n), 0.0, 0.01, 100, tv=tv)
    for prec, tol in (("f64", 1e-12), ("f32", 2e-7)):
        with T.Engine(n, prec, "rk4", "mf") as e:
            e.set_option(T.OPT_TILED, 2)
            e.set_time_variant(tv)
            e.set_x0_uniform()
            e.run(0.0, 0.01, 100)
            x = e.get_x()
        assert np.max(np.abs(x - ox)) <= tol, prec


def _ref_run(n, r, steps, x0=None, tv=None, **opts):
    with T.Engine(n, "f64", "ref", "mf") as e:
        for k, v in opts.items():
            e.set_option(getattr(T, "OPT_" + k.upper()), int(v))
        if tv is not None:
            e.set_time_variant(tv)
        else:
            e.set_rates(r)
        if x0 is None:
            e.set_x0_uniform()
        else:
            e.set_x(x0)
        info = e.run(0.0, 0.01, steps, patience=-1)
        return e.get_x(), info


@pytest.mark.parametrize("n", [13, 17, 20])
def test_tiled_ref_bit_exact_vs_per_state_kernel(n):
    """The tiled reference loop keeps the reference's term order: bitwise
    equal to the one-lane-per-state kernel, any tile order or grid."""
    rng = np.random.default_rng(700 + n)
    r = rng.uniform(0.1, 2.0, n + 1)
    x0 = rng.random(1 << n)
    x0 /= x0.sum()
    a, ia = _ref_run(n, r, 30, x0, tiled=2)
    b, ib = _ref_run(n, r, 30, x0, tiled=0)
    c, _ = _ref_run(n, r, 30, x0, tiled=2, xcd_order=0, tile_blocks=64)
    d, _ = _ref_run(n, r, 30, x0, tiled=2, xcd_order=4)
    assert _bits_equal(a, b) and _bits_equal(a, c) and _bits_equal(a, d)
    assert ia.t == ib.t and abs(ia.last_checksum - ib.last_checksum) <= 1e-14 * ib.last_checksum


def test_tiled_ref_time_variant_bit_exact_vs_per_state_kernel():
    n = 14
    rng = np.random.default_rng(71)
    tv = dict(type=(np.arange(n + 1) % 3).astype(np.int32), amp=rng.uniform(0, 0.5, n + 1),
              period=rng.uniform(0.5, 2, n + 1), offset=rng.uniform(0, 1, n + 1), c=np.ones(n + 1))
    a, _ = _ref_run(n, None, 40, tv=tv, tiled=2)
    b, _ = _ref_run(n, None, 40, tv=tv, tiled=0)
    assert _bits_equal(a, b)


def test_graph_replay_with_level_walk_n22():
    """n=22 FP32 RK4 and the FP64 reference loop on tiles in the level walk:
    the walk's table is built before the first captured chunk (16 steps at
    n=22), and 40 steps = 2 replays + 8 direct steps equal direct launches
    bit for bit."""
    n = 22
    r = np.random.default_rng(222).uniform(0.2, 1.5, n + 1)
    x0 = _random_state(n, 223)
    for prec, integ in (("f32", "rk4"), ("f64", "ref")):
        out = []
        for graph in (1, 0):
            with T.Engine(n, prec, integ, "mf") as e:
                e.set_option(T.OPT_GRAPH, graph)
                e.set_rates(r)
                e.set_x(x0)
                info = e.run(0.0, 0.01, 40, patience=-1)
                assert info.steps_done == 40
                out.append(e.get_x())
        assert _bits_equal(out[0], out[1]), (prec, integ)


def test_graph_replay_does_not_change_results():
    """n=18: chunks of 256 steps are replayed from a captured graph; 600 steps
    = 2 replays + 88 direct steps, bitwise equal to direct launches."""
    n = 18
    rng = np.random.default_rng(81)
    r = rng.uniform(0.2, 1.5, n + 1)
    tv = dict(type=np.zeros(n + 1, np.int32), amp=rng.uniform(0, 0.5, n + 1),
              period=rng.uniform(0.5, 2, n + 1), offset=np.zeros(n + 1), c=np.ones(n + 1))
    for prec, integ, kind in (("f64", "ref", "r"), ("f32", "rk4", "r"), ("f32", "rk4", "tv"),
                              ("f64", "ref", "tv")):
        out = []
        for graph in (1, 0):
            with T.Engine(n, prec, integ, "mf") as e:
                e.set_option(T.OPT_GRAPH, graph)
                if kind == "tv":
                    e.set_time_variant(tv)
                else:
                    e.set_rates(r)
                e.set_x0_state(0)
                info = e.run(0.0, 0.01, 600, patience=-1)
                out.append((e.get_x(), info))
        assert np.array_equal(out[0][0], out[1][0]), (prec, integ, kind)
        assert out[0][1].t == out[1][1].t and out[0][1].steps_done == out[1][1].steps_done == 600
    # early exit inside a replayed chunk: alpha = beta = 1/2, gamma = 1, uniform
    with T.Engine(n, "f64", "ref", "mf") as e:
        e.set_rates(np.array([0.5] + [1.0] * (n - 1) + [0.5]))
        e.set_x0_uniform()
        info = e.run(0.0, 0.01, 2000, thr=1e-7, patience=100)
    assert info.stopped and info.steps_at_stop == 100 and info.steps_done == 101


# ------------------------------------------- LDS-resident whole-run kernel
def _engine_run(n, prec, integ, r=None, tv=None, x0=None, steps=100, h=0.01, thr=1e-7,
                patience=-1, pieces=1, **opts):
    with T.Engine(n, prec, integ, "mf") as e:
        for k, v in opts.items():
            e.set_option(getattr(T, "OPT_" + k.upper()), int(v))
        if tv is not None:
            e.set_time_variant(tv)
        else:
            e.set_rates(r)
        if x0 is None:
            e.set_x0_state(0)
        else:
            e.set_x(x0)
        per = steps // pieces
        for i in range(pieces):
            k = per if i < pieces - 1 else steps - per * (pieces - 1)
            info = e.run(0.0, h, k, thr=thr, patience=patience, resume=i > 0)
            if info.stopped:
                break
        return e.get_x(), info


def _tv(n, seed):
    rng = np.random.default_rng(seed)
    return dict(type=(np.arange(n + 1) % 3).astype(np.int32), amp=rng.uniform(0, 0.5, n + 1),
                period=rng.uniform(0.5, 2, n + 1), offset=rng.uniform(0, 1, n + 1),
                c=np.ones(n + 1))


@pytest.mark.parametrize("n", [2, 3, 6, 9, 11, 13])
def test_lds_ref_bit_exact_vs_oracle_and_streaming(n):
    r = np.random.default_rng(200 + n).uniform(0.1, 2.0, n + 1)
    x0 = np.random.default_rng(300 + n).random(1 << n)
    x0 /= x0.sum()
    ox, oinfo = oracle.run_ref(n, r, x0, 0.0, 0.01, 120, patience=-1)
    xl, il = _engine_run(n, "f64", "ref", r, x0=x0, steps=120, lds=2)
    xs, is_ = _engine_run(n, "f64", "ref", r, x0=x0, steps=120, lds=0)
    assert _bits_equal(xl, ox), f"max|dx| = {np.max(np.abs(xl - ox))}"
    assert _bits_equal(xs, ox)
    assert il.t == oinfo.t and il.steps_done == 120
    assert abs(il.last_checksum - oinfo.last_checksum) <= 1e-14 * oinfo.last_checksum
    assert abs(il.mass - 1.0) <= 1e-12


@pytest.mark.parametrize("threads", [64, 128, 512, 1024])
def test_lds_workgroup_size_does_not_change_results(threads):
    n = 12
    r = np.random.default_rng(9).uniform(0.1, 2.0, n + 1)
    base, _ = _engine_run(n, "f64", "ref", r, steps=80, lds=0)
    x, _ = _engine_run(n, "f64", "ref", r, steps=80, lds=2, lds_threads=threads)
    assert _bits_equal(x, base)


def test_lds_ref_early_stop_and_chunks_match_streaming():
    """n=10 homogeneous rates (alpha = beta = 1/2) from the empty lattice:
    10000 steps (early exit armed, not reached) in one call and in 7 resumed
    pieces, LDS vs streaming."""
    n = 10
    r = np.array([0.5] + [1.0] * (n - 1) + [0.5])
    h = oracle.step_h(100.0, 0.0, 10000)
    a, ia = _engine_run(n, "f64", "ref", r, steps=10000, h=h, patience=100, lds=1)
    b, ib = _engine_run(n, "f64", "ref", r, steps=10000, h=h, patience=100, lds=1, pieces=7)
    c, ic = _engine_run(n, "f64", "ref", r, steps=10000, h=h, patience=100, lds=0)
    assert _bits_equal(a, c) and _bits_equal(b, c)
    assert ia.steps_done == ib.steps_done == ic.steps_done
    assert ia.t == ib.t == ic.t
    # early exit from the stationary uniform state: leaves at index 100
    u, iu = _engine_run(n, "f64", "ref", r, x0=oracle.uniform_x0(n), steps=10000, h=h,
                        patience=100, lds=1)
    assert iu.stopped and iu.steps_at_stop == 100 and iu.steps_done == 101


def test_lds_csr_reference_loop_matches_streaming():
    """The stored-CSR reference loop inside one workgroup: bitwise equal to
    the one-launch-per-step CSR kernel and to the matrix-free runs."""
    n = 11
    r = np.random.default_rng(33).uniform(0.3, 1.5, n + 1)
    h = oracle.step_h(50.0, 0.0, 5000)
    out = []
    for gen, lds in (("csr", 1), ("csr", 0), ("mf", 1)):
        with T.Engine(n, "f64", "ref", gen) as e:
            e.set_option(T.OPT_LDS, lds)
            e.set_rates(r)
            e.set_x0_state(0)
            info = e.run(0.0, h, 5000, thr=1e-7, patience=100)
            out.append((e.get_x(), info))
    for x, info in out[1:]:
        assert _bits_equal(x, out[0][0])
        assert info.steps_done == out[0][1].steps_done and info.t == out[0][1].t


@pytest.mark.parametrize("gen", ["mf", "csr"])
def test_windowed_exit_test_matches_oracle(gen):
    """One lane per state with the exit test armed: steps run in windows in
    which no exit is possible, then the reductions replay the exit logic
    step by step; the stop step, t and x equal the oracle's reference loop."""
    n = 13
    r = np.random.default_rng(41).uniform(0.5, 1.5, n + 1)
    h = 0.05
    ox, oinfo = oracle.run_ref(n, r, oracle.uniform_x0(n), 0.0, h, 40000, thr=1e-7, patience=100,
                               use_csr=gen == "csr")
    assert oinfo.stopped
    with T.Engine(n, "f64", "ref", gen) as e:
        e.set_option(T.OPT_LDS, 0)
        e.set_rates(r)
        e.set_x0_uniform()
        info = e.run(0.0, h, 40000, thr=1e-7, patience=100)
        x = e.get_x()
    assert info.stopped and info.steps_at_stop == oinfo.steps_at_stop
    assert info.steps_done == oinfo.steps_done and info.t == oinfo.t
    assert _bits_equal(x, ox)


def test_lds_converged_exit_matches_streaming():
    n = 8
    r = np.random.default_rng(31).uniform(0.3, 1.5, n + 1)
    a, ia = _engine_run(n, "f64", "ref", r, steps=100000, h=0.02, thr=1e-7, patience=100, lds=1)
    b, ib = _engine_run(n, "f64", "ref", r, steps=100000, h=0.02, thr=1e-7, patience=100, lds=0)
    assert ia.stopped and ib.stopped
    assert ia.steps_at_stop == ib.steps_at_stop and ia.steps_done == ib.steps_done
    assert _bits_equal(a, b)


@pytest.mark.parametrize("n", [4, 8, 12, 13])
def test_lds_rk4_vs_oracle(n):
    r = np.random.default_rng(400 + n).uniform(0.2, 1.5, n + 1)
    x0 = np.zeros(1 << n)
    x0[0] = 1.0
    ox, _ = oracle.run_rk4(n, r, x0, 0.0, 0.01, 150)
    x64, _ = _engine_run(n, "f64", "rk4", r, steps=150, lds=2)
    x32, _ = _engine_run(n, "f32", "rk4", r, steps=150, lds=2)
    assert np.max(np.abs(x64 - ox)) <= 1e-12
    err = np.max(np.abs(x32 - ox))
    print(f"n={n} FP32 LDS max|x - x_f64| = {err:.3e}")
    assert err <= 1e-6
    s64, _ = _engine_run(n, "f64", "rk4", r, steps=150, lds=2, rk4_staged=1)
    assert np.max(np.abs(s64 - x64)) <= 1e-14


@pytest.mark.parametrize("n", [7, 13])
def test_lds_time_variant_vs_oracle(n):
    tv = _tv(n, 50 + n)
    x0 = oracle.uniform_x0(n)
    ox, _ = oracle.run_rk4(n, None, x0, 0.0, 0.01, 120, tv=tv)
    for prec, tol in (("f64", 1e-12), ("f32", 2e-7)):
        x, info = _engine_run(n, prec, "rk4", tv=tv, x0=x0, steps=120, patience=100, lds=2)
        assert info.steps_done == 120 and not info.stopped
        assert np.max(np.abs(x - ox)) <= tol, prec
    # reference loop with time-variant rates: LDS and streaming agree bitwise
    a, _ = _engine_run(n, "f64", "ref", tv=tv, x0=x0, steps=120, lds=2)
    b, _ = _engine_run(n, "f64", "ref", tv=tv, x0=x0, steps=120, lds=0)
    assert _bits_equal(a, b)


# ------------------------------------------------------------ error handling
def test_errors_are_reported():
    with pytest.raises(T.TasepError):
        T.Engine(40)
    with pytest.raises(T.TasepError):
        T.Engine(10, "f32", "ref", "mf")
    with T.Engine(6, "f64", "ref", "mf") as e:
        with pytest.raises(T.TasepError):
            e.run(0.0, 0.01, 10)          # no rates yet
        with pytest.raises(T.TasepError):
            e.set_rates(np.array([0.5, -1, 1, 1, 1, 1, 0.5]))
        with pytest.raises(T.TasepError):
            e.set_x0_state(1 << 6)
        with pytest.raises(T.TasepError):
            e.set_option(T.OPT_LDS_THREADS, 96)   # not a power of two
        with pytest.raises(T.TasepError):
            e.set_option(T.OPT_LDS, 3)
    with T.Engine(3, "f64", "ref", "mf") as e:
        with pytest.raises(T.TasepError):
            e.window_marginals()          # needs n >= 4


@pytest.mark.parametrize("n,opts", [(10, dict(lds=1)), (16, dict(lds=0, tiled=0)), (20, dict(lds=0, tiled=1)),
                                    (16, dict(lds=0, tiled=2))])
def test_ref_loop_zero_rates_bit_exact_vs_oracle(n, opts):
    """Rates of 0 (no entry, no exit, blocked bonds): every active term adds
    0 * x = +0, which must leave each row sum as the reference's, on the
    LDS, one-lane and tiled kernels."""
    rng = np.random.default_rng(80 + n)
    for zero in ((0,), (n,), (3, n - 4), (0, n)):
        r = rng.uniform(0.2, 1.8, n + 1)
        r[list(zero)] = 0.0
        x0 = _product_np(n, _site_densities(n, 81 + n))
        ox, oinfo = oracle.run_ref(n, r, x0, 0.0, 0.01, 30, patience=-1, use_csr=False)
        with T.Engine(n, "f64", "ref", "mf") as e:
            for k, v in opts.items():
                e.set_option(getattr(T, "OPT_" + k.upper()), v)
            e.set_rates(r)
            e.set_x(x0)
            info = e.run(0.0, 0.01, 30, patience=-1)
            x = e.get_x()
        assert info.steps_done == 30
        assert _bits_equal(x, ox), (zero, float(np.max(np.abs(x - ox))))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_rk4_zero_rates_vs_oracle(prec, tol):
    """Classical RK4 with blocked bonds and no exit: the mass stays 1 and
    the state follows the oracle, relative to its largest entry (tiles at
    n=18 and one lane at n=15)."""
    for n in (15, 18):
        r = np.random.default_rng(90 + n).uniform(0.2, 1.8, n + 1)
        r[[2, n - 5, n]] = 0.0
        x0 = oracle.uniform_x0(n)
        ox, _ = oracle.run_rk4(n, r, x0, 0.0, 0.01, 25)
        with T.Engine(n, prec, "rk4", "mf") as e:
            e.set_option(T.OPT_LDS, 0)
            e.set_rates(r)
            e.set_x0_uniform()
            info = e.run(0.0, 0.01, 25, patience=-1)
            x = e.get_x()
        rel = float(np.max(np.abs(x - ox)) / np.max(ox))
        print(f"n={n} {prec}: max|x - x_oracle| / max x = {rel:.2e}")
        assert rel <= tol
        assert abs(info.mass - 1.0) <= (1e-12 if prec == "f64" else 1e-5)
