"""The reference update divides by 6 with IEEE division
(src/masterEquation.c:640).  The device kernels replace that division by an
FMA-corrected product (div6 in tasep_amd/csrc/generator.hpp): q = RN(a*y),
y = RN(1/6), r = fma(-6, q, a), q' = copysign(fma(r, y, q), a).  This test
restates that sequence in C on the CPU (same IEEE binary64 operations) and
checks it against a / 6.0, bit for bit, on random operands over the whole
exponent range, on operands next to products 6*q (near-midpoint quotients)
and on signed zeros and extremes.  The GPU parity tests then pin the kernels
themselves bitwise against the oracle."""
import os
import shutil
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double q6(double a) {
    const double y = 1.0 / 6.0;
    const double m = fabs(a);
    if ((m < 0x1p-960 || m > 0x1p1000) && a != 0.0) return a / 6.0;
    double q = a * y;
    double r = fma(-6.0, q, a);
    return copysign(fma(r, y, q), a);
}
static long check(double a) {
    double t = a / 6.0, u = q6(a);
    return memcmp(&t, &u, 8) != 0;
}
int main(void) {
    long bad = 0, n = 0;
    for (long i = 0; i < 20000000L; i++, n++) {
        uint64_t bits = (xr() & 0x800FFFFFFFFFFFFFull) | ((uint64_t)(1 + xr() % 2046) << 52);
        double a; memcpy(&a, &bits, 8);
        bad += check(a);
    }
    for (long i = 0; i < 5000000L; i++) {
        uint64_t bits = (xr() & 0x000FFFFFFFFFFFFFull) | ((uint64_t)(1023 + (long)(xr() % 64) - 32) << 52);
        double q; memcpy(&q, &bits, 8);
        double a = 6.0 * q;
        for (int k = -3; k <= 3; k++, n++) {
            double b = a;
            for (int j = 0; j < (k < 0 ? -k : k); j++) b = nextafter(b, k < 0 ? -INFINITY : INFINITY);
            bad += check(b);
        }
    }
    bad += check(0.0) + check(-0.0) + check(0x1p-1074) + check(0x1p-960) + check(0x1.fffffffffffffp+1023)
         + check(INFINITY) + check(-INFINITY) + check(0x1p1000) + check(-0x1.8p999);
    printf("%ld %ld\n", n, bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_fma_corrected_division_by_6_is_correctly_rounded():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "div6.c"), os.path.join(d, "div6")
        with open(src, "w") as f:
            f.write(SRC)
        # -ffp-contract=off: only the explicit fma() calls fuse
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-mfma", src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    n, bad = int(out[0]), int(out[1])
    assert n > 50_000_000 and bad == 0
