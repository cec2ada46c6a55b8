"""The fast integrator's sincos (gargantua_amd/csrc/fast_math.hpp) against
double-precision sin and cos on the CPU. The function uses only fmaf and
single-rounded operations, so this host build rounds exactly like gfx950."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "fast_math.hpp"
static double ulp_err(float got, double want) {
    float w = (float)want;
    float next = nextafterf(fabsf(w), INFINITY);
    double ulp = (double)next - (double)fabsf(w);
    if (w == 0.0f) ulp = 1.4e-45;
    return fabs((double)got - want) / ulp;
}
int main(void) {
    double worst_s = 0, worst_c = 0, worst_abs = 0, worst_x = 0;
    for (long i = -4000000; i <= 4000000; ++i) {
        float x = (float)(i * 1e-5);   /* [-40, 40] */
        float s, c;
        garg::fast_sincos(x, &s, &c);
        double ws = sin((double)x), wc = cos((double)x);
        double es = ulp_err(s, ws), ec = ulp_err(c, wc);
        double ea = fmax(fabs(s - ws), fabs(c - wc));
        if (fabs(ws) > 1e-3 && es > worst_s) { worst_s = es; worst_x = x; }
        if (fabs(wc) > 1e-3 && ec > worst_c) worst_c = ec;
        if (ea > worst_abs) worst_abs = ea;
    }
    /* small angles near the pole keep relative accuracy */
    double worst_small = 0;
    for (int k = 1; k < 100000; ++k) {
        float x = (float)k * 1e-7f;
        float s, c;
        garg::fast_sincos(x, &s, &c);
        worst_small = fmax(worst_small, fabs(s - sin((double)x)) / sin((double)x));
    }
    printf("%g %g %g %g %g\n", worst_s, worst_c, worst_abs, worst_small, worst_x);
    return 0;
}
"""


def test_fast_sincos_accuracy():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "h.cpp"), os.path.join(d, "h")
        open(src, "w").write(HARNESS)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "gargantua_amd", "csrc"), src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    worst_s, worst_c, worst_abs, worst_small = (float(v) for v in out[:4])
    print(f"sin {worst_s:.2f} ulp, cos {worst_c:.2f} ulp, abs {worst_abs:.2e}, small-angle rel {worst_small:.2e}")
    assert worst_s < 4 and worst_c < 4 and worst_abs < 3e-7 and worst_small < 2e-7


SHIFT = r"""
#include <math.h>
#include <stdio.h>
#include "fast_math.hpp"
int main(void) {
    double worst = 0, worst_fallback = 0;
    for (int i = 0; i <= 4000; ++i) {
        float x0 = -1.0f + 5.0f * (float)i / 4000.0f;   /* theta near [0, pi] and past the poles */
        float s0, c0;
        garg::fast_sincos(x0, &s0, &c0);
        for (int j = -500; j <= 500; ++j) {
            float d = 0.3f * (float)j / 500.0f;         /* inside and past the series limit */
            float x = x0 + d;
            float s, c;
            garg::sincos_shift(s0, c0, d, x, &s, &c);
            double want_s = sin((double)x0 + (double)d), want_c = cos((double)x0 + (double)d);
            double e = fmax(fabs(s - want_s), fabs(c - want_c));
            if (fabsf(d) <= GARG_SHIFT_LIMIT) worst = fmax(worst, e);
            else worst_fallback = fmax(worst_fallback, fmax(fabs(s - sin((double)x)), fabs(c - cos((double)x))));
        }
    }
    printf("%g %g\n", worst, worst_fallback);
    return 0;
}
"""


def test_sincos_shift_accuracy():
    """Angle addition from a known sine and cosine: absolute error within a
    few float ulp of 1 over the shift range, fast_sincos beyond it."""
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "s.cpp"), os.path.join(d, "s")
        open(src, "w").write(SHIFT)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "gargantua_amd", "csrc"), src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    worst, fallback = float(out[0]), float(out[1])
    print(f"shifted sincos abs error {worst:.2e}, fallback {fallback:.2e}")
    assert worst < 4e-7 and fallback < 3e-7



VERSINE = r"""
#include <math.h>
#include <stdio.h>
#include "fast_math.hpp"
int main(void) {
    /* relative error against 2 sin^2(d / 2) in double, which has no
       cancellation: tiny steps, the series range and the fallback */
    double worst_series = 0, worst_fallback = 0;
    for (long i = 1; i <= 2000000; ++i) {
        float d = (float)(i * 2.5e-7);           /* (0, 0.5] */
        double h = sin((double)d * 0.5);
        double want = 2.0 * h * h;
        double err = fabs((double)garg::versine(d) - want) / want;
        double err_neg = fabs((double)garg::versine(-d) - want) / want;
        if (err_neg > err) err = err_neg;
        if (d <= 0.25f) { if (err > worst_series) worst_series = err; }
        else if (err > worst_fallback) worst_fallback = err;
    }
    for (int k = -15; k <= -3; ++k) {             /* down to 1e-15 rad (d^2 stays a normal float) */
        float d = powf(10.0f, (float)k);
        double h = sin((double)d * 0.5);
        double err = fabs((double)garg::versine(d) - 2.0 * h * h) / (2.0 * h * h);
        if (err > worst_series) worst_series = err;
    }
    printf("%.3e %.3e\n", worst_series, worst_fallback);
    return 0;
}
"""


def test_versine_relative_accuracy():
    """1 - cos d for the Kerr chord: full relative accuracy for tiny steps
    (where float 1 - cos d is zero) through the series range, and the
    fallback beyond it, against 2 sin^2(d / 2) in double."""
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "v.cpp"), os.path.join(d, "v")
        open(src, "w").write(VERSINE)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "gargantua_amd", "csrc"), src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    series, fallback = float(out[0]), float(out[1])
    print(f"versine relative error: series {series:.2e}, fallback {fallback:.2e}")
    assert series < 6e-7 and fallback < 2e-6


SMALL = r"""
#include <math.h>
#include <stdio.h>
#include "fast_math.hpp"
int main(void) {
    double worst_atan = 0, worst_log = 0, worst_sin = 0;
    for (int i = -200000; i <= 200000; ++i) {
        if (i == 0) continue;
        float x = 0.2f * (float)i / 200000.0f;      /* inside and past each series limit */
        worst_atan = fmax(worst_atan, fabs(garg::small_atan(x) - atan((double)x)) / fabs(atan((double)x)));
        worst_log = fmax(worst_log, fabs(garg::small_log1p(x) - log1p((double)x)) / fabs(log1p((double)x)));
        float d = 1.5f * x;
        worst_sin = fmax(worst_sin, fabs(garg::small_sin(d) - sin((double)d)) / fabs(sin((double)d)));
    }
    /* tiny arguments keep relative accuracy */
    for (int k = 1; k < 1000; ++k) {
        float x = (float)k * 1e-9f;
        worst_atan = fmax(worst_atan, fabs(garg::small_atan(x) - atan((double)x)) / atan((double)x));
        worst_log = fmax(worst_log, fabs(garg::small_log1p(x) - log1p((double)x)) / log1p((double)x));
        worst_sin = fmax(worst_sin, fabs(garg::small_sin(x) - sin((double)x)) / sin((double)x));
    }
    printf("%g %g %g\n", worst_atan, worst_log, worst_sin);
    return 0;
}
"""


def test_small_argument_series():
    """small_atan, small_log1p and small_sin (the Kerr-Schild frame
    rotation's step-to-step differences) within a few float roundings of
    the double-precision functions, relative to the result, across and
    past each series' limit and at tiny arguments."""
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "h.cpp"), os.path.join(d, "h")
        open(src, "w").write(SMALL)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "gargantua_amd", "csrc"), src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    worst_atan, worst_log, worst_sin = (float(v) for v in out)
    print(f"relative error: atan {worst_atan:.2e}, log1p {worst_log:.2e}, sin {worst_sin:.2e}")
    assert worst_atan < 3e-7 and worst_log < 3e-7 and worst_sin < 3e-7
