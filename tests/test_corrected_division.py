"""The smoothers' corrected-reciprocal division equals IEEE division where they use it.

The HIP smoothers divide by 6 (Poisson) and by d = 1 + 6 alpha (diffusion) as
q1 = x * RN(1/d), q = fma(fma(-q1, d, x), RN(1/d), q1), sign copied from x
(stable-fluids_amd/csrc/sf_relax.hpp: div6, div_by; the stage quotients of the black-cell
and full-iterate passes use the float64 form of test_float64_quotient.py), and fall back to the
division instruction outside 2^-100 <= |x| <= 2^100 (2^-110 and FLT_MAX for d = 6) or
for d outside [1, 2^20]. Markstein's theorem makes the corrected quotient the correctly
rounded one when RN(1/d) is correctly rounded and nothing under- or overflows; this test
checks that claim on the host (C, IEEE single precision with a correctly rounded fmaf)
against x / d over every exponent of the guarded range (a fixed stride through each
binade) for divisors that include all-ones mantissas, and that zeros keep their sign.
"""
import os
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float corrected(float x, float d, float r) {
    const float q1 = x * r;
    const float e = fmaf(-q1, d, x);
    return copysignf(fmaf(e, r, q1), x);
}
int main(void) {
    const float divisors[] = {6.0f, 1.0f, 1.0000006f, 1.00006f, 1.39321601f, 1.5f, 1.99999988f, 2.0f, 3.0f,
                              3.99999976f, 7.0f, 12.3456789f, 1000.5f, 65535.996f, 131073.0f, 1048575.9f};
    long checked = 0, bad = 0;
    for (unsigned k = 0; k < sizeof(divisors) / sizeof(divisors[0]); ++k) {
        const float d = divisors[k], r = 1.0f / d;
        const uint32_t lo = d == 6.0f ? 0x08800000u : 0x0d800000u;  /* 2^-110 or 2^-100 */
        const uint32_t hi = d == 6.0f ? 0x7f7fffffu : 0x71800000u;  /* FLT_MAX or 2^100 */
        for (uint32_t b = lo; b <= hi && b >= lo; b += 97u) {
            for (int s = 0; s < 2; ++s) {
                const float x = f_of(b | (s ? 0x80000000u : 0u));
                ++checked;
                if (u_of(corrected(x, d, r)) != u_of(x / d)) {
                    if (bad < 5) printf("mismatch d=%a x=%a\n", d, x);
                    ++bad;
                }
            }
        }
        for (int s = 0; s < 2; ++s) {  /* zeros: exact via the copied sign */
            const float z = s ? -0.0f : 0.0f;
            ++checked;
            if (u_of(corrected(z, d, r)) != u_of(z / d)) ++bad;
        }
    }
    printf("%ld %ld\n", checked, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_corrected_reciprocal_equals_ieee_division(tmp_path):
    c = tmp_path / "div.c"
    exe = tmp_path / "div"
    c.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", str(c), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    checked, bad = (int(v) for v in out.stdout.strip().split("\n")[-1].split())
    assert out.returncode == 0 and bad == 0, out.stdout
    assert checked > 50_000_000
