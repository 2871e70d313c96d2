"""The kernels divide by the constants 3 and 6 as ``q = a * RN(1/b)``,
``r = fma(-q, b, a)``, ``RN(q + r RN(1/b))`` (``div_const`` in
``xlb-biot_amd/csrc/xb_device.h``), which Markstein's theorem makes the correctly
rounded quotient for finite, non-underflowing ``a``: bit-identical to the division
the oracle performs.  Checked here on the CPU (IEEE fma, as the GPU's
v_fma_f64 / v_fma_f32) against the division on random FP64 operands over exponents
-600 .. 600 and on every FP32 significand at several exponents."""
import os
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 88172645463325252ull;
static uint64_t rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double q64(double a, double b, double inv) { double q = a * inv; double r = fma(-q, b, a); return fma(r, inv, q); }
static float q32(float a, float b, float inv) { float q = a * inv; float r = fmaf(-q, b, a); return fmaf(r, inv, q); }
int main(void) {
  long bad = 0, n = 0;
  const double bs[2] = {3.0, 6.0};
  for (int k = 0; k < 2; ++k) {
    const double b = bs[k], inv = 1.0 / b;
    for (long i = 0; i < 40000000L; ++i) {
      uint64_t bits = (rnd() & 0x800FFFFFFFFFFFFFull) | ((uint64_t)(1023 - 600 + rnd() % 1201) << 52);
      double a, x, y;
      memcpy(&a, &bits, 8);
      x = a / b; y = q64(a, b, inv); ++n;
      if (memcmp(&x, &y, 8) != 0 && !(x == 0 && y == 0)) ++bad;
    }
    const float bf = (float)b, invf = 1.0f / bf;
    for (int ex = -60; ex <= 60; ex += 20)
      for (uint32_t m = 0; m < (1u << 23); ++m) {
        uint32_t bits = ((uint32_t)(127 + ex) << 23) | m;
        float a, x, y;
        memcpy(&a, &bits, 4);
        x = a / bf; y = q32(a, bf, invf); ++n;
        if (memcmp(&x, &y, 4) != 0) ++bad;
      }
  }
  printf("%ld %ld\n", n, bad);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_division_by_constant_is_correctly_rounded(tmp_path):
    c = tmp_path / "div.c"
    c.write_text(SRC)
    exe = tmp_path / "div"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"], check=True)
    n, bad = map(int, subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split())
    assert n > 1.5e8 and bad == 0, (n, bad)
