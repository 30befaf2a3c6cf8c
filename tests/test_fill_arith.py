"""The fill kernel's quotient offset / w, checked over its whole domain.

The fill kernel divides the window offset by the fill width with the rounded
reciprocal r = RN(1/w) and one residual step (q0 = RN(o r), e = o - q0 w exact
by fma, q = RN(q0 + e r)) instead of an IEEE division per pixel. The oracle
divides (fill_area.inc.glsl:21, t = offset / (right.x - left.x)). Packed fill
coordinates are multiples of 1/256 in [0, 256), so w = k / 256 for k in
[1, 65535], and for a pixel column the fill touches, offset = O / 512 with
O in [0, 2k]. This test runs every such pair through both and requires the
same float.
"""
import os
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
int main(int argc, char **argv) {
    int wmax = atoi(argv[1]);
    long long bad = 0, total = 0;
    #pragma omp parallel for reduction(+:bad,total) schedule(dynamic, 64)
    for (int k = 1; k <= wmax; k++) {
        volatile float wv = (float)k / 256.0f;
        float w = wv;
        float r = 1.0f / w;
        for (int o2 = 0; o2 <= 2 * k; o2++) {
            float o = (float)o2 / 512.0f;
            float q = o / w;
            float q0 = o * r;
            float e = fmaf(-q0, w, o);
            float q1 = fmaf(e, r, q0);
            total++;
            bad += q1 != q;
        }
    }
    printf("%lld %lld\n", total, bad);
    return 0;
}
"""


def _has_fma():
    try:
        with open("/proc/cpuinfo") as f:
            return " fma " in f.read()
    except OSError:
        return False


def test_reciprocal_division_matches_ieee_division(tmp_path):
    if not _has_fma():
        pytest.skip("host has no hardware fma (the software fmaf would take minutes)")
    src = tmp_path / "div.c"
    src.write_text(SRC)
    exe = tmp_path / "div"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-mfma", "-ffp-contract=off", "-o", str(exe),
                    str(src), "-lm"], check=True)
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    out = subprocess.run([str(exe), "65535"], check=True, capture_output=True, text=True,
                         env=env).stdout.split()
    total, bad = int(out[0]), int(out[1])
    assert total == sum(2 * k + 1 for k in range(1, 65536))
    assert bad == 0


def test_planted_wrong_step_is_caught(tmp_path):
    """The same check with the residual step left out finds mismatches."""
    if not _has_fma():
        pytest.skip("host has no hardware fma")
    src = tmp_path / "div_bad.c"
    src.write_text(SRC.replace("float q1 = fmaf(e, r, q0);", "float q1 = q0; (void)e;"))
    exe = tmp_path / "div_bad"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-mfma", "-ffp-contract=off", "-o", str(exe),
                    str(src), "-lm"], check=True)
    out = subprocess.run([str(exe), "4096"], check=True, capture_output=True, text=True).stdout
    assert int(out.split()[1]) > 0


BYTE_SRC = r"""
#include <math.h>
#include <stdio.h>
int main(void) {
    int bad = 0;
    const float r = 1.0f / 255.0f;
    for (int c = 0; c < 256; c++) {
        volatile float xv = (float)c;
        float x = xv;
        float q0 = x * r;
        float q = fmaf(fmaf(-q0, 255.0f, x), r, q0);
        bad += q != x / 255.0f;
    }
    printf("%d\n", bad);
    return 0;
}
"""


def test_byte_over_255_from_reciprocal(tmp_path):
    """The fill kernel turns clip-mask bytes into floats with the same residual
    step (fill.cs.glsl:84 loads the clip texel as byte / 255); all 256 bytes
    must give the IEEE quotient."""
    if not _has_fma():
        pytest.skip("host has no hardware fma")
    src = tmp_path / "byte.c"
    src.write_text(BYTE_SRC)
    exe = tmp_path / "byte"
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"],
                   check=True)
    assert subprocess.run([str(exe)], check=True, capture_output=True,
                          text=True).stdout.strip() == "0"
