"""Exact steady state of the open TASEP (matrix-product ansatz), exact rationals.

TEST INFRASTRUCTURE ONLY (checker for converged runs of the HIP path).

Restates the reference's GMP oracle ``src/steadystates_gmp.c:20-186`` with
``fractions.Fraction`` instead of 64-bit ``mpf``:

* ``B(n, p) = p (2n-p-1)! / (n! (n-p)!)``           (``eval_B_gmp``, :20-65)
* ``Z_0 = 1``, ``Z_{i+1} = sum_{p=0}^{i+1} sum_{q=0}^{p} a^q b^(p-q) B(i+1,p)``
  with ``a = 1/alpha``, ``b = 1/beta`` and ``B(., 0) = 0``   (``allocfill_Z``, :97-140)
* ``rho_i`` (i = 0 is the entry site), ``i < n-1``:
  ``[sum_{k=0}^{n-i-2} B(k+1,1) Z_{n-1-k} + b^(k+2) B(n-i-1,k+1) Z_i] / Z_n``;
  ``rho_{n-1} = Z_{n-1} / (beta Z_n)``                        (``allocfill_Rho``, :142-186)

The model it solves has bulk hop rate 1; a homogeneous bulk rate gamma is the
same chain with time rescaled, so ``(alpha, beta, gamma)`` has the densities
of ``(alpha/gamma, beta/gamma, 1)``.  Inputs are taken as the exact binary
values of the given doubles (the reference rounds 1/alpha to a double first,
:118-119; this module keeps 1/alpha exact).

Two-point function (``two_point_function``, :188-326), sites 1-based from the
entry, i < j < N:

* ``<tau_i tau_j>_N = sum_{p=0}^{N-j-1} C_p (Z_{N-1-p}/Z_N) <tau_i>_{N-1-p}
  + (Z_{j-1}/Z_N) <tau_i>_{j-1} sum_{p=2}^{N-j+1} (p-1)(2(N-j)-p)! /
  ((N-j)! (N-j+1-p)!) beta^-p``, C_p = (2p)!/(p!(p+1)!) (``two_point``);
  j = N, which the reference excludes (``assert(j<N)``, :191), from the
  boundary vector: ``<tau_i tau_N>_N = Z_{N-1} <tau_i>_{N-1} / (beta Z_N)``.
* The reference's own sums take every Z through ``calculate_Z_i(m)``
  (:329-340), which returns ``Z[m-1]`` of an array whose entry m is Z_m, so
  each Z in its result is one index lower than in the formula above.
  ``two_point_reference`` restates that literally; it equals the exact pair
  moment only where the Z are geometric (alpha + beta = 1), and it is what
  ``steadystates_gmp -i -j`` prints.

``steady_state_bruteforce`` solves ``L p = 0`` exactly over Fractions for
small n from the generator definition alone; tests use it to pin this module.
"""
from __future__ import annotations

from fractions import Fraction
from math import factorial


def _F(v) -> Fraction:
    return v if isinstance(v, Fraction) else Fraction(v)


def B(n: int, p: int) -> Fraction:
    if p < 1 or p > n:
        return Fraction(0)
    return Fraction(p * factorial(2 * n - p - 1), factorial(n) * factorial(n - p))


def Z_list(n: int, alpha, beta):
    a, b = 1 / _F(alpha), 1 / _F(beta)
    Z = [Fraction(1)]
    for i in range(n):
        acc = Fraction(0)
        for p in range(0, i + 2):
            bp = B(i + 1, p)
            if bp == 0:
                continue
            for q in range(0, p + 1):
                acc += a ** q * b ** (p - q) * bp
        Z.append(acc)
    return Z


def densities(n: int, alpha, beta, gamma=1):
    """Exact <tau_i>, i = 0 entry site ... n-1 exit site, as Fractions."""
    g = _F(gamma)
    al, be = _F(alpha) / g, _F(beta) / g
    Z = Z_list(n, al, be)
    b = 1 / be
    rho = []
    for i in range(n - 1):
        acc = Fraction(0)
        for k in range(0, n - i - 1):
            acc += B(k + 1, 1) * Z[n - 1 - k]
            acc += b ** (k + 2) * B(n - i - 1, k + 1) * Z[i]
        rho.append(acc / Z[n])
    rho.append(Z[n - 1] / (be * Z[n]))
    return rho


def _catalan(p: int) -> Fraction:
    return Fraction(factorial(2 * p), factorial(p) * factorial(p + 1))


def _boundary_sum(N: int, j: int, b: Fraction) -> Fraction:
    """sum_{p=2}^{N-j+1} (p-1)(2(N-j)-p)! / ((N-j)!(N-j+1-p)!) b^p, b = 1/beta
    (the second sum of steadystates_gmp.c:283-318)."""
    m = N - j
    acc = Fraction(0)
    for p in range(2, m + 2):
        acc += Fraction((p - 1) * factorial(2 * m - p), factorial(m) * factorial(m + 1 - p)) * b ** p
    return acc


def _two_point(i: int, j: int, alpha, beta, N: int, gamma, zshift: int) -> Fraction:
    g = _F(gamma)
    al, be = _F(alpha) / g, _F(beta) / g
    if not (1 <= i < j <= N):
        raise ValueError("need 1 <= i < j <= N")
    Zs = Z_list(N, al, be)
    Z = lambda m: Zs[m - zshift]  # noqa: E731
    if j == N and zshift == 0:
        return Zs[N - 1] * densities(N - 1, al, be)[i - 1] / (be * Zs[N])
    if j == N:
        raise ValueError("the reference's two-point function needs j < N")
    first = Fraction(0)
    for p in range(0, N - j):                        # :230-279
        first += _catalan(p) * Z(N - 1 - p) / Z(N) * densities(N - p - 1, al, be)[i - 1]
    second = _boundary_sum(N, j, 1 / be) * Z(j - 1) * densities(j - 1, al, be)[i - 1] / Z(N)
    return first + second


def two_point(i: int, j: int, alpha, beta, N: int, gamma=1) -> Fraction:
    """Exact stationary <tau_i tau_j> (sites 1-based from the entry,
    1 <= i < j <= N), as a Fraction."""
    return _two_point(i, j, alpha, beta, N, gamma, 0)


def two_point_reference(i: int, j: int, alpha, beta, N: int) -> Fraction:
    """The value steadystates_gmp's two_point_function computes (exact
    arithmetic in place of 64-bit mpf), with its Z index one lower than the
    formula's (calculate_Z_i, :329-340); 1 <= i < j < N."""
    return _two_point(i, j, alpha, beta, N, 1, 1)


def current(n: int, alpha, beta, gamma=1):
    """Stationary current J = alpha (1 - rho_entry) (= beta rho_exit)."""
    rho = densities(n, alpha, beta, gamma)
    return _F(alpha) * (1 - rho[0])


def steady_state_bruteforce(n: int, rates):
    """Exact stationary distribution p (index = state bitmask) from L p = 0,
    sum p = 1, with rates r[0]=alpha, r[1..n-1] bulk, r[n]=beta (Fractions).

    Independent of the matrix-product formulas: builds the generator from the
    hop rules (bit n-1 entry site, bit 0 exit site) and eliminates exactly.
    """
    r = [_F(v) for v in rates]
    dim = 1 << n
    # column-oriented out-rates: L[t][s] for s -> t
    L = [[Fraction(0)] * dim for _ in range(dim)]
    for s in range(dim):
        out = []
        if not (s >> (n - 1)) & 1:
            out.append((s | (1 << (n - 1)), r[0]))
        if s & 1:
            out.append((s & ~1, r[n]))
        for k in range(n - 1):
            if ((s >> (k + 1)) & 1) and not ((s >> k) & 1):
                out.append((s - (1 << (k + 1)) + (1 << k), r[n - 1 - k]))
        for t, w in out:
            L[t][s] += w
            L[s][s] -= w
    # replace last equation with normalisation
    A = [row[:] for row in L]
    A[dim - 1] = [Fraction(1)] * dim
    rhs = [Fraction(0)] * dim
    rhs[dim - 1] = Fraction(1)
    # Gauss-Jordan
    for c in range(dim):
        piv = next(i for i in range(c, dim) if A[i][c] != 0)
        A[c], A[piv] = A[piv], A[c]
        rhs[c], rhs[piv] = rhs[piv], rhs[c]
        inv = 1 / A[c][c]
        A[c] = [v * inv for v in A[c]]
        rhs[c] *= inv
        for i in range(dim):
            if i != c and A[i][c] != 0:
                f = A[i][c]
                A[i] = [vi - f * vc for vi, vc in zip(A[i], A[c])]
                rhs[i] -= f * rhs[c]
    return rhs


def densities_from_p(n: int, p):
    """<tau> entry..exit from a distribution indexed by bitmask."""
    return [sum(p[s] for s in range(1 << n) if (s >> (n - 1 - j)) & 1)
            for j in range(n)]
