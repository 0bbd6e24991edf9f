"""Division by a constant divisor (sxh_div_markstein in sx_ops.h).

The specialised kernels divide by a tape constant c as q = x * RN(1/c) with
two corrections q += (x - c q) * RN(1/c) through exact FMA remainders; the
second gives the correctly rounded quotient by Markstein's theorem (a
faithful quotient and a reciprocal within half an ulp).  These tests run the
same header function on the host against IEEE division, bit for bit:

* near-midpoint dividends, built exactly: for an odd T in [2^53, 2^54) with
  C * T = X * 2^j + r (C the divisor's 53-bit significand, |r| < 2^12 small),
  x = X scaled has x / c = T / 2 (a midpoint between two doubles) minus
  r / (2C) ulp -- within about 2^-41 ulp of a rounding boundary, the cases
  a wrong rounding shows in first.  Without the corrections (q = x * RN(1/c))
  a third of them come out one ulp off, so the set has teeth;
* random significands over the exponent range the constant-divisor path takes
  (2^-900 <= |x| <= 2^900, 2^-64 <= |c| <= 2^64).
"""
import math
import os
import random
import shutil
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SX_OPS = os.path.join(HERE, "..", "casadi_amd", "csrc", "sx_ops.h")

DIVISORS = [1.1, 3.0, 0.1, 7.0, 1e-3, math.pi, 0.3, 9.81, 1.0 / 3.0, 2.5, math.sqrt(0.5), 1.0000000000000002,
            1.9999999999999998, 1.2, 0.05, 10.0, 12.0, 0.7, 1.7, -1.1, -0.3, 2.0 ** -64, 2.0 ** 64, 1.5e19]


def near_midpoint_dividends(c, count, rng):
    """Dividends x whose quotient x / c lies within ~2^-41 ulp of a midpoint."""
    m, e = math.frexp(abs(c))
    C = int(m * 2 ** 53)                   # c = C * 2^(e - 53), C in [2^52, 2^53)
    t = (C & -C).bit_length() - 1
    codd = C >> t
    out = []
    while len(out) < count:
        r = rng.randrange(1, 1 << 12) | 1
        sgn = rng.choice((1, -1))
        for j in (53, 54):
            mod = 1 << (j - t)
            if mod <= 1:
                continue
            t0 = (sgn * r * pow(codd, -1, mod)) % mod
            if t0 % 2 == 0:
                continue
            lo, hi = -((t0 - (1 << 53)) // mod), ((1 << 54) - 1 - t0) // mod
            if lo > hi:
                continue
            T = t0 + rng.randint(lo, hi) * mod          # odd, C * T = r * 2^t mod 2^j
            P = C * T
            if P.bit_length() - 53 != j:
                continue
            X = (P - sgn * r * (1 << t)) >> j
            if not (1 << 52) <= X < (1 << 53):
                continue
            x = math.ldexp(X, rng.randrange(-800, 800))
            out.append(-x if rng.random() < 0.5 else x)
            break
    return out


def _random_dividends(count, rng):
    out = []
    for _ in range(count):
        x = math.ldexp(1.0 + rng.getrandbits(52) * 2.0 ** -52, rng.randrange(-900, 900))
        out.append(-x if rng.random() < 0.5 else x)
    return out


DRIVER = r"""
#include <cstdint>
#include <cstdio>
#include <cstring>
#include "sx_ops.h"
// argv[1]: binary file of (divisor, count, dividends...) records; argv[2]:
// corrections (2: the shipped function, 0: the bare product x * RN(1/c)).
int main(int argc, char** argv) {
  FILE* f = std::fopen(argv[1], "rb");
  const int steps = argv[2][0] - '0';
  double c;
  long long n;
  long long bad = 0, tot = 0;
  while (std::fread(&c, 8, 1, f) == 1) {
    if (std::fread(&n, 8, 1, f) != 1) return 2;
    for (long long i = 0; i < n; ++i) {
      double x;
      if (std::fread(&x, 8, 1, f) != 1) return 2;
      const double ci = 1.0 / c;
      const double q = steps == 2 ? sxh::sxh_div_markstein(x, c, ci) : x * ci;
      const double r = x / c;
      bad += std::memcmp(&q, &r, 8) != 0;
      ++tot;
    }
  }
  std::printf("%lld %lld\n", tot, bad);
  return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("div")
    src = d / "div.cpp"
    src.write_text(DRIVER)
    exe = d / "div"
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.dirname(SX_OPS), str(src),
                    "-o", str(exe)], check=True)
    return str(exe), d


def _run(driver, records, steps):
    exe, d = driver
    path = d / f"cases_{steps}.bin"
    with open(path, "wb") as f:
        for c, xs in records:
            f.write(struct.pack("<dq", c, len(xs)) + np.asarray(xs, dtype="<f8").tobytes())
    out = subprocess.run([exe, str(path), str(steps)], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), int(out[1])


def test_near_midpoint_quotients_correctly_rounded(driver):
    rng = random.Random(20261203)
    records = [(c, near_midpoint_dividends(c, 20000, rng)) for c in DIVISORS]
    tot, bad = _run(driver, records, 2)
    assert tot == 20000 * len(DIVISORS) and bad == 0
    # the cases are hard: the bare reciprocal product misrounds many of them
    _, bad0 = _run(driver, records, 0)
    assert bad0 > tot // 10


def test_random_quotients_correctly_rounded(driver):
    rng = random.Random(7)
    records = [(c, _random_dividends(50000, rng)) for c in DIVISORS]
    tot, bad = _run(driver, records, 2)
    assert tot == 50000 * len(DIVISORS) and bad == 0
