"""Host logic of the bulk kernels' tile decomposition: the multiply-high and
shift pairs that replace the division of a workgroup's tile index by the
tiles per row, the tile rows per plane and the rows per plane
(fast_div_params in mplb_amd/csrc/lattice_kernels.hpp, used by k_bulk and
k_lattice).  A small C++ harness includes the header itself and checks the
quotients against integer division: every n below 5,000, the multiples of d
and their neighbours, the top of the 31-bit range, and random n, for every
d up to 5,000, the powers of two and their neighbours, and random d."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cstdint>
#include <cstdio>
#include <random>
#include "lattice_kernels.hpp"
static int fd(int n, uint32_t m, int sh) { return (int)(((uint64_t)(uint32_t)n * m) >> sh); }
int main() {
    std::mt19937_64 r(7);
    long bad = 0, checks = 0;
    auto test = [&](int32_t d) {
        uint32_t m; int32_t sh;
        mplbhip::fast_div_params(d, &m, &sh);
        auto chk = [&](int64_t n) {
            if (n < 0 || n >= (1LL << 31)) return;
            checks++;
            if (fd((int)n, m, sh) != (int)(n / d) && bad++ < 5)
                printf("wrong quotient d=%d n=%lld\n", d, (long long)n);
        };
        for (int64_t n = 0; n < 5000; n++) chk(n);
        for (int64_t k = 1; k < 2000; k++) { chk(k * d - 1); chk(k * d); chk(k * d + 1); }
        for (int64_t n = (1LL << 31) - 2000; n < (1LL << 31); n++) chk(n);
        const int64_t top = ((1LL << 31) - 1) / d * d;
        chk(top); chk(top - 1);
        for (int i = 0; i < 2000; i++) chk((int64_t)(r() % (1ULL << 31)));
    };
    for (int32_t d = 1; d <= 5000; d++) test(d);
    for (int i = 0; i < 3000; i++) test((int32_t)(1 + r() % ((1ULL << 31) - 1)));
    for (int l = 0; l < 31; l++) { test(1 << l); test((1 << l) + 1); if (l > 1) test((1 << l) - 1); }
    test(2147483647);
    printf("checks %ld wrong %ld\n", checks, bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fast_div_quotients_are_exact(tmp_path):
    src = tmp_path / "fast_div.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "fast_div"
    subprocess.run(["g++", "-O2", "-std=c++17",
                    "-I", os.path.join(ROOT, "mplb_amd", "csrc"),
                    "-D__host__=", "-D__device__=",
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "wrong 0" in out.stdout
