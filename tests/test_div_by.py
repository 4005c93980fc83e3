"""The kernels divide by the cell width dx as div_by(a, dx, 1/dx): q0 = a * r,
then q0 + fma(-q0, dx, a) * r with fmas (Markstein's correction). Bitwise
parity with the oracle's plain `(x - xmin) / dx` rests on that being the
correctly rounded quotient. This test checks it against IEEE division in C
(the same FP64 operations, with fma from libm) on random quotients: positions
across the domain, values near 0, arbitrary exponents, and near-multiples of
dx (quotients next to an integer, where trunc() picks the cell). A second
generator draws from the region where the error bound of q0 is loosest: the
divisor's significand just below 2 (so 1/dx's significand is just above 1
and its half-ulp error is largest relative to the quotient) and the
quotient's significand just below 2 as well.
"""
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static uint64_t s = 88172645463325252ull;
static uint64_t xs(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double u01(void) { return (xs() >> 11) * 0x1.0p-53; }
int main(int argc, char **argv) {
    const long per = atol(argv[1]);
    double dxs[16] = {(2 * M_PI / 1.25) / 32, (4 * M_PI) / 32, (4 * M_PI) / 16,
                      (2 * M_PI / 1.5) / 64, 1.0 / 3, 0.1, 0.7, 4 * M_PI / 128};
    for (int d = 8; d < 16; ++d) dxs[d] = (2 * M_PI / (1.0 + 0.5 * u01())) / (16 << (xs() % 3));
    long bad = 0;
    for (int d = 0; d < 16; ++d) {
        const double dx = dxs[d], r = 1.0 / dx;
        for (long i = 0; i < per; ++i) {
            double a;
            const int k = (int)(xs() % 4);
            if (k == 0) a = u01() * 64 * dx;
            else if (k == 1) a = (u01() - 0.5) * 1e-3;
            else if (k == 2) {
                uint64_t b = xs() & 0x800fffffffffffffull;
                b |= (uint64_t)(1023 - 60 + xs() % 70) << 52;
                memcpy(&a, &b, 8);
            } else a = dx * (double)(xs() % 200) + ((xs() & 1) ? 1 : -1) * ldexp(u01(), -50);
            const double q0 = a * r;
            const double q = fma(fma(-q0, dx, a), r, q0);
            if (q != a / dx) ++bad;
        }
    }
    printf("%ld\n", bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_by_is_the_correctly_rounded_quotient(tmp_path):
    src = tmp_path / "div_by.c"
    exe = tmp_path / "div_by"
    src.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "2000000"], check=True, capture_output=True, text=True)
    assert int(out.stdout.strip()) == 0


SRC_HARD = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t xs(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double u01(void) { return (xs() >> 11) * 0x1.0p-53; }
int main(int argc, char **argv) {
    const long n = atol(argv[1]);
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        /* divisor: significand in [2 - 2^-8, 2), any exponent in [-8, 8] */
        const double bsig = 2.0 - ldexp(u01(), -8);
        const double b = ldexp(bsig, (int)(xs() % 17) - 8);
        /* quotient target: significand in [2 - 2^-8, 2) */
        const double t = ldexp(2.0 - ldexp(u01(), -8), (int)(xs() % 9) - 4);
        double a = t * b;
        if (xs() & 1) { uint64_t u; memcpy(&u, &a, 8); u += (xs() % 7) - 3; memcpy(&a, &u, 8); }
        const double r = 1.0 / b;
        const double q0 = a * r;
        const double q = fma(fma(-q0, b, a), r, q0);
        if (q != a / b) ++bad;
    }
    printf("%ld\n", bad);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_by_in_the_loosest_region(tmp_path):
    src = tmp_path / "div_hard.c"
    exe = tmp_path / "div_hard"
    src.write_text(SRC_HARD)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "16000000"], check=True, capture_output=True, text=True)
    assert int(out.stdout.strip()) == 0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_by_check_detects_a_plain_reciprocal_product(tmp_path):
    """The same harness with the correction step dropped (q = a * (1/dx))
    must report mismatches, so the check above can fail."""
    src = tmp_path / "recip.c"
    exe = tmp_path / "recip"
    src.write_text(SRC.replace("const double q = fma(fma(-q0, dx, a), r, q0);",
                               "const double q = q0;"))
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    out = subprocess.run([str(exe), "200000"], check=True, capture_output=True, text=True)
    assert int(out.stdout.strip()) > 0
