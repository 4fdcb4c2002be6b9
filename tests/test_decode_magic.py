"""Index decode by invariant division (``gm_magic`` in ``amg_internal.h``).

The kernels map flat chequer indices to (i, j, k) as
``q = (t + ((n - t) >> s1)) >> s2`` with ``t = mulhi(m, n)`` and the
multiplier and shifts that ``gm_magic(d)`` computes on the host (Granlund and
Montgomery, "Division by invariant integers using multiplication", 1994).
This test compiles the header's ``gm_magic`` into a host program and checks
the quotient against exact integer division for every divisor up to 1100,
over every n below 2^16, the top 20 000 32-bit values, 20 000 random values
and every multiple of d (and its neighbours) below 2^20.  Levels have
hx = nx/2 and ny far below that bound; the GPU parity tests cover the
kernels that decode.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"

SRC = r"""
#include "amg_internal.h"
#include <cstdint>
#include <cstdio>
static uint32_t q(uint32_t n, uint32_t m, int sh) {
  const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
  return (t + ((n - t) >> (sh & 0xff))) >> (sh >> 8);
}
int main() {
  unsigned long long bad = 0, tested = 0;
  for (uint32_t d = 1; d <= 1100; ++d) {
    int sh = 0;
    const uint32_t m = amg::gm_magic(d, &sh);
    auto chk = [&](uint32_t n) { ++tested; if (q(n, m, sh) != n / d) ++bad; };
    for (uint32_t n = 0; n < (1u << 16); ++n) chk(n);
    for (uint32_t n = 0xFFFFFFFFu - 20000u; n != 0; ++n) chk(n);
    uint64_t x = 0x9E3779B97F4A7C15ull ^ d;
    for (int r = 0; r < 20000; ++r) { x ^= x << 13; x ^= x >> 7; x ^= x << 17; chk((uint32_t)x); }
    for (uint64_t n = d; n < (1u << 20); n += d) { chk((uint32_t)n - 1); chk((uint32_t)n); chk((uint32_t)n + 1); }
  }
  std::printf("%llu %llu\n", tested, bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not found")
def test_invariant_division_is_exact(tmp_path):
    src = tmp_path / "gm.cpp"
    src.write_text(SRC)
    exe = tmp_path / "gm"
    r = subprocess.run([HIPCC, "-O2", "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "anelastic_solve_amd", "csrc"), str(src), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    tested, bad = (int(x) for x in r.stdout.split())
    assert r.returncode == 0 and bad == 0 and tested > 10**8, r.stdout
