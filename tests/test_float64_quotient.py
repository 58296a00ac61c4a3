"""The black-cell smoothers' float64 quotient equals IEEE single-precision division
wherever they use it.

StageDiv (stable-fluids_amd/csrc/sf_relax.hpp) divides by 6 (Poisson) and by d = 1 + 6 alpha
(diffusion) as q = RN32(RN64(x * RN64(1/d))), and sends a stage through the division
instruction when some numerator is nonzero with |x| < d * 2^-126 (a subnormal quotient),
detected through the key (bits(x) << 1) - 2. The claim: for every float x that is not sent
back, q equals x / d. This test checks it on the host (C, IEEE binary64 and binary32
arithmetic) for every sign and a fixed stride through all finite floats, every float of
the lowest binades, zeros and infinities, for divisors that include all-ones mantissas,
and that the key flags exactly the nonzero |x| below d * 2^-126.
"""
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
static uint32_t key(float x) { return (u_of(x) << 1) - 2u; }
static long checked = 0, bad = 0, flag_bad = 0;
static void check(float x, float d, double rd, uint32_t tkey, float t) {
    const int flagged = key(x) < tkey;
    const int expect  = fabsf(x) < t && x != 0.0f;   /* nonzero |x| below d * 2^-126 */
    ++checked;
    if (flagged != expect) {
        if (flag_bad < 5) printf("flag mismatch d=%a x=%a\n", d, x);
        ++flag_bad;
    }
    if (flagged) return;                             /* the exact path takes it */
    const float q = (float)((double)x * rd);
    const float e = x / d;
    if (u_of(q) != u_of(e) && !(isnan(q) && isnan(e))) {
        if (bad < 5) printf("mismatch d=%a x=%a q=%a x/d=%a\n", d, x, q, e);
        ++bad;
    }
}
int main(void) {
    const float divisors[] = {6.0f, 1.0f, 1.0000006f, 1.00006f, 1.39321601f, 1.5f, 1.99999988f, 2.0f, 3.0f,
                              3.99999976f, 4.93216f, 7.0f, 12.3456789f, 1000.5f, 65535.996f, 131073.0f, 1048575.9f};
    for (unsigned k = 0; k < sizeof(divisors) / sizeof(divisors[0]); ++k) {
        const float d   = divisors[k];
        const double rd = 1.0 / (double)d;
        const float t   = d * 0x1p-126f;                /* exact: d >= 1 */
        const uint32_t tkey = key(t);
        for (uint32_t b = 0; b <= 0x7f7fffffu; b += 97u)
            for (int s = 0; s < 2; ++s) check(f_of(b | (s ? 0x80000000u : 0u)), d, rd, tkey, t);
        for (uint32_t b = 0; b < 0x02000000u; ++b)      /* every float below 2^-122 */
            for (int s = 0; s < 2; ++s) check(f_of(b | (s ? 0x80000000u : 0u)), d, rd, tkey, t);
        check(INFINITY, d, rd, tkey, t);
        check(-INFINITY, d, rd, tkey, t);
        check(0x1.fffffep127f, d, rd, tkey, t);
        check(-0x1.fffffep127f, d, rd, tkey, t);
    }
    printf("%ld %ld %ld\n", checked, bad, flag_bad);
    return bad != 0 || flag_bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_float64_quotient_equals_ieee_division(tmp_path):
    c = tmp_path / "q64.c"
    exe = tmp_path / "q64"
    c.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", str(c), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    checked, bad, flag_bad = (int(v) for v in out.stdout.strip().split("\n")[-1].split())
    assert out.returncode == 0 and bad == 0 and flag_bad == 0, out.stdout
    assert checked > 1_000_000_000
