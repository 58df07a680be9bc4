"""The snow step divides every sub-cell storage by the cell's land area fraction
with the reciprocal computed once per cell: q0 = x (1/b), r = fma(-q0, b, x),
q = fma(r, 1/b, q0).  For a correctly rounded reciprocal this gives the correctly
rounded quotient (Markstein's theorem); this test compares it with the IEEE
division on random operands, including divisors whose significands are all ones
and quotients of products s * b / b, as the kernel forms them.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
static uint64_t st[2] = {0x9E3779B97F4A7C15ull, 0xD1B54A32D192ED03ull};
static uint64_t rnd(void) { uint64_t a = st[0], b = st[1]; st[0] = b; a ^= a << 23; st[1] = a ^ b ^ (a >> 17) ^ (b >> 26); return st[1] + b; }
static double u01(void) { return (rnd() >> 11) * 0x1.0p-53; }
int main(int argc, char **argv) {
  const long n = atol(argv[1]);
  long bad = 0;
  for (long i = 0; i < n; i++) {
    double b, x;
    const int mode = rnd() % 4;
    if (mode == 0) b = u01();
    else if (mode == 1) b = 1.0 - u01() * 1e-6;
    else if (mode == 2) {
      uint64_t m = rnd() & ((1ull << 52) - 1);
      if (rnd() % 2) m = ((1ull << 52) - 1) ^ (rnd() & 0xff);
      const uint64_t bits = (1022ull << 52) | m;
      memcpy(&b, &bits, 8);
    } else b = ldexp(1.0, -(int)(rnd() % 20)) * (1.0 + u01()) * 100.0;
    if (!(b > 0)) continue;
    const double y = 1.0 / b;
    const int xm = rnd() % 3;
    if (xm == 0) x = u01() * 1000.0 * b;
    else if (xm == 1) x = ldexp(1.0 + u01(), (int)(rnd() % 60) - 40);
    else x = (u01() * 1e4) * b;
    const double q = x / b;
    const double q0 = x * y;
    const double m = fma(fma(-q0, b, x), y, q0);
    if (memcmp(&q, &m, 8) != 0) bad++;
  }
  printf("%ld %ld\n", n, bad);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_reciprocal_division_is_correctly_rounded():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "mdiv.c"), os.path.join(d, "mdiv")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe, "20000000"], check=True, capture_output=True, text=True).stdout.split()
    assert int(out[0]) == 20000000 and int(out[1]) == 0
