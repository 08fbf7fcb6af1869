"""The fast quotient of the temporally blocked stepper, checked on the CPU.

The stepper may replace ``x / d`` (the reference's ``tensor / dx``,
src/core/fdtd2d.py:232-235 and :244-257) by a reciprocal product whose
residual is corrected at the scale s = 2^k that puts ``d*s`` in [2^31, 2^32)::

    q0 = x*r;  t = fma(q0, d*s, -(x*s));  q = fma(-t, r/s, q0)

It is used for every ``|x| < 2^(96+e)`` (``d = m*2^e``), the range that
``engine.division_range`` hands to the device's exhaustive check. This test
compiles a small C program (glibc's correctly rounded ``fmaf``) that compares
the form with IEEE division, bit for bit, on every float of the binades where
a residual step goes wrong first: all subnormals, the lowest normal binades,
the binades around 2^-103 where an unscaled residual stops being exact, and
the top binades of the range; plus a seeded sample of the rest. The same
comparison must catch the two plausible mistakes: the residual taken at scale
1, and no correction step.
"""

import subprocess
import shutil

import numpy as np
import pytest

from waveforge_amd.engine import division_range

CHECK_C = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float D, R, D2, RK, S;
static int uncorrected;
static long bad;
static void one(uint32_t u) {
    for (int sg = 0; sg < 2; ++sg) {
        volatile float x = fb(u | (sg ? 0x80000000u : 0u));
        float q0 = x * R, x2 = x * S;
        float t = fmaf(q0, D2, -x2);
        float q = uncorrected ? q0 : fmaf(-t, RK, q0);
        volatile float ref = x / D;
        if (bf(q) != bf(ref)) ++bad;
    }
}
static void band(uint32_t lo, uint32_t hi) { for (uint32_t u = lo; u < hi; ++u) one(u); }
int main(int argc, char **argv) {
    D = strtof(argv[1], 0);
    R = 1.0f / D;
    uint32_t span = (uint32_t)strtoul(argv[2], 0, 10);
    uint32_t db = bf(D), k = 158u - ((db >> 23) & 0xffu);
    if (argc > 3 && !strcmp(argv[3], "unscaled")) k = 0;  /* residual at scale 1 */
    uncorrected = argc > 3 && !strcmp(argv[3], "uncorrected");
    D2 = fb(db + (k << 23));
    RK = fb(bf(R) - (k << 23));
    S = fb((k + 127u) << 23);
    band(0u, 3u << 23);                              /* subnormals, 2^-126, 2^-125 */
    band((127u - 106u) << 23, (127u - 101u) << 23);  /* 2^-106 .. 2^-102 */
    band(span - (2u << 23), span);                   /* the top two binades */
    uint64_t st = 88172645463325252ull;
    for (int n = 0; n < 4000000; ++n) {              /* seeded sample of the rest */
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        one((uint32_t)(st % span));
    }
    printf("%ld\n", bad);
    return 0;
}
"""


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    d = tmp_path_factory.mktemp("fastdiv")
    (d / "check.c").write_text(CHECK_C)
    exe = d / "check"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", str(exe), str(d / "check.c"), "-lm"], check=True)
    return exe


def _run_all(exe, jobs):
    """Run the checker on (divisor, mistake) jobs side by side; mismatch counts."""
    procs = []
    for dval, mistake in jobs:
        span = division_range(dval)[1]
        args = [str(exe), repr(float(np.float32(dval))), str(span)] + ([mistake] if mistake else [])
        procs.append(subprocess.Popen(args, stdout=subprocess.PIPE, text=True))
    return [int(p.communicate(timeout=300)[0].strip()) for p in procs]


DIVISORS = [1e-3, 5e-4, 2.5e-4, 1.5625e-3, 1e-2, 3.7, 1e-6]


def test_fast_quotient_equals_ieee(checker):
    assert _run_all(checker, [(d, None) for d in DIVISORS]) == [0] * len(DIVISORS)


def test_planted_mistakes_are_caught(checker):
    assert all(n > 0 for n in _run_all(checker, [(1e-3, "unscaled"), (1e-3, "uncorrected")]))
