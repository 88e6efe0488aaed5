"""randint_fast (the samplers' division-free u % n, csrc/philox.h) against
the plain 64-bit modulo of the reference's randint (random.cc:174-176):
boundary dividends and divisors plus random pairs, bit-exact.  The header
is compiled for the host with g++ (the same source the device code uses)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "mcmc_for_ammsb_amd", "csrc")

PROGRAM = r"""
#include <stdio.h>
#include "philox.h"
using namespace ammsb;
static uint64_t sm = 0x9E3779B97F4A7C15ull;
static uint64_t next() {  // splitmix64
  uint64_t z = (sm += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static long bad = 0, checked = 0;
static void check(uint64_t u, uint32_t n) {
  const uint32_t want = (uint32_t)(u % n), got = randint_fast(u, n, fastmod_recip(n));
  ++checked;
  if (want != got && bad++ < 5)
    printf("mismatch u=%llu n=%u want %u got %u\n", (unsigned long long)u, n, want, got);
}
int main() {
  const uint32_t ns[] = {1u, 2u, 3u, 5u, 7u, 1000u, 5000u, 65536u, 1000000u, 4000000u,
                         50120564u, 161619936u, 2147483647u, 2147483648u, 4294967291u,
                         4294967295u};
  for (uint32_t n : ns) {
    const uint64_t us[] = {0ull, 1ull, n - 1ull, (uint64_t)n, n + 1ull, 2ull * n - 1,
                           ~0ull, ~0ull - n, ~0ull - n + 1, (~0ull / n) * n,
                           (~0ull / n) * n - 1, 1ull << 63, (1ull << 32) - 1};
    for (uint64_t u : us) check(u, n);
    for (int i = 0; i < 20000; ++i) check(next(), n);
  }
  for (int i = 0; i < 2000000; ++i) {
    uint32_t n = (uint32_t)next();
    if (n == 0) n = 1;
    if (i & 1) n >>= (next() % 31);
    if (n == 0) n = 1;
    check(next(), n);
  }
  printf("checked %ld bad %ld\n", checked, bad);
  return bad ? 1 : 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_randint_fast_equals_modulo(tmp_path):
    src = tmp_path / "fastmod.cc"
    src.write_text(PROGRAM)
    exe = tmp_path / "fastmod"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert " bad 0" in out.stdout
