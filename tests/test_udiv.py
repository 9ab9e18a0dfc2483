"""The multiply-shift division of the kernels' tile indexing (csrc/eup_udiv.h), compiled with the
host C++ compiler and checked against integer division (CPU).

Every divisor up to 20 000 against numerators at, next to and between its multiples and the top
of the range, then 10^6 random (numerator, divisor) pairs over the whole range n < 2^31,
1 <= d < 2^31."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "euplectella_hpc_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include <random>
#include "eup_udiv.h"
int main() {
    std::mt19937_64 g(20261202);
    long bad = 0;
    auto check = [&](uint32_t n, uint32_t d, const eup::UDiv &u) {
        if (n >= 0x80000000u) return;
        if (u.div(n) != n / d && bad++ < 5) std::printf("d=%u n=%u got %u\n", d, n, u.div(n));
    };
    for (uint32_t d = 1; d <= 20000; ++d) {
        const eup::UDiv u = eup::UDiv::make(d);
        check(0x7FFFFFFFu, d, u);
        for (int r = 0; r < 50; ++r) {
            const uint32_t n = (uint32_t)(g() & 0x7FFFFFFFu);
            const uint64_t q = n / d;
            check(n, d, u);
            check((uint32_t)(q * d), d, u);
            if (q) check((uint32_t)(q * d - 1), d, u);
            check((uint32_t)(q * d + d - 1), d, u);
        }
    }
    for (int r = 0; r < 1000000; ++r) {
        const uint32_t d = (uint32_t)(g() % 0x7FFFFFFFu) + 1u;
        check((uint32_t)(g() & 0x7FFFFFFFu), d, eup::UDiv::make(d));
    }
    std::printf("mismatches %ld\n", bad);
    return bad != 0;
}
"""


def test_multiply_shift_division_is_exact(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "udiv_check.cpp"
    exe = tmp_path / "udiv_check"
    src.write_text(HARNESS)
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout
    assert "mismatches 0" in res.stdout
