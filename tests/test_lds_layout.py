"""The tile pass's LDS regions (program.h lds_layout, used by both the kernel and the host
launcher) are disjoint, aligned and within the 160 KiB of a gfx950 CU for every tile size the
scheduler can pick.  Compiled with the host compiler; no GPU needed."""
import os
import subprocess
import tempfile

import pytest

from tests.conftest import ROOT

SRC = r"""
#include <cstdio>
#include <initializer_list>
#include "program.h"
int main() {
    for (int amp : {8, 16})
        for (int adj = 0; adj < 2; adj++)
            for (int K = mqa::kMinTileQubits; K <= mqa::kMaxTileQubits; K++)
                for (int C = 0; C <= K; C++) {
                    auto l = mqa::lds_layout(K, C, adj, amp, mqa::kMaxTapsPerPass);
                    std::printf("%d %d %d %d %u %u %u %u %u\n", amp, adj, K, C, l.psi, l.lam, l.offtab, l.taps, l.total);
                }
    return 0;
}
"""


def test_lds_regions_disjoint():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "lay.cpp")
        exe = os.path.join(d, "lay")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "mindquantum_amd", "csrc"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows = [list(map(int, line.split())) for line in out.strip().splitlines()]
    assert rows
    for amp, adj, K, C, psi, lam, offtab, taps, total in rows:
        tile = (1 << K) * amp
        regions = [(psi, psi + tile)]
        if adj:
            regions.append((lam, lam + tile))
        regions.append((offtab, offtab + (1 << (K - C)) * 8))
        regions.append((taps, total))
        for (a0, a1), (b0, b1) in zip(regions, regions[1:]):
            assert a1 <= b0, (amp, adj, K, C)
        assert offtab % 8 == 0 and taps % 16 == 0
        assert lam == tile  # lambda follows psi


@pytest.mark.parametrize("amp,adj,K,C", [(8, 0, 12, 5), (8, 1, 12, 5), (16, 0, 11, 4), (16, 1, 11, 4)])
def test_default_tiles_fit_one_cu(amp, adj, K, C):
    tile = (1 << K) * amp * (2 if adj else 1)
    assert tile + (1 << (K - C)) * 8 + 96 * ((1 << (K - 4)) // 64 or 1) * 16 <= 160 * 1024


SWZ = r"""
#include <cstdio>
#include <vector>
#include "program.h"
int main() {
    int bad = 0;
    for (int amp : {8, 16}) {
        for (int K = mqa::kMinTileQubits; K <= mqa::kMaxTileQubits; K++) {
            const unsigned n = 1u << K;
            std::vector<int> seen(n, 0);
            for (unsigned u = 0; u < n; u++) {
                unsigned s = mqa::lds_swizzle(u, amp);
                if (s >= n) bad++; else seen[s]++;
            }
            for (unsigned u = 0; u < n; u++) if (seen[u] != 1) bad++;
            for (unsigned a = 0; a < n; a += 37)
                for (unsigned b = 0; b < n; b += 53)
                    if (mqa::lds_swizzle(a ^ b, amp) != (mqa::lds_swizzle(a, amp) ^ mqa::lds_swizzle(b, amp))) bad++;
            if (amp == 8) {
                // complex64 tile I/O moves amplitudes u, u + 1 (u even) through the 16-byte slot at
                // swz(u) & ~1, halves swapped when swz(u) is odd; every amplitude must land on swz(u)
                std::vector<unsigned> lds(n, ~0u);
                for (unsigned u = 0; u < n; u += 2) {
                    if (mqa::lds_swizzle(u | 1, amp) != (mqa::lds_swizzle(u, amp) ^ 1u)) bad++;
                    const unsigned a = mqa::lds_swizzle(u, amp), slot = a & ~1u;
                    lds[slot] = (a & 1u) ? u + 1 : u;
                    lds[slot + 1] = (a & 1u) ? u : u + 1;
                }
                for (unsigned u = 0; u < n; u++) if (lds[mqa::lds_swizzle(u, amp)] != u) bad++;
            }
        }
    }
    std::printf("%d\n", bad);
    return 0;
}
"""


def test_lds_swizzle_is_a_linear_bijection():
    """The kernel forms LDS slots as swz(base) ^ swz(offset); that needs XOR-linearity, and the
    slots must be a permutation of the tile for every tile size.  The complex64 tile I/O moves
    amplitude pairs through 16-byte slots, which needs swz(u | 1) == swz(u) ^ 1."""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "swz.cpp")
        exe = os.path.join(d, "swz")
        with open(src, "w") as f:
            f.write(SWZ.replace("#include <cstdio>", "#include <cstdio>\n#include <initializer_list>"))
        subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "mindquantum_amd", "csrc"), src, "-o", exe],
                       check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == 0
