"""The driver's number formatter (cpu-vh_amd/csrc/vh_format.hpp) against
Python's %-formatting, which rounds the exact binary value correctly with ties
to even, as glibc's printf does: "%g" (surface.dat, an ostream's default
output, freezeout/FreezeOut.cpp:243-270) and "%.3f"/"%.8f" (the snapshot
files, rhic/src/FileIO.cpp:17,46-56)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <cstdio>
#include <string>
#include <vector>
#include "vh_format.hpp"
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb");
    std::vector<double> v;
    double x;
    while (std::fread(&x, sizeof x, 1, f) == 1) v.push_back(x);
    std::fclose(f);
    std::string out;
    for (double y : v) {
        cpuvh::put_g(out, y);
        out += ' ';
        cpuvh::put_fixed(out, y, 3);
        out += ' ';
        cpuvh::put_fixed(out, y, 8);
        out += '\n';
    }
    std::fwrite(out.data(), 1, out.size(), stdout);
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fmt")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cpu-vh_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    return exe, d


def _values():
    rng = np.random.default_rng(7)
    n = 400_000
    wide = rng.uniform(1, 10, n) * 10.0 ** rng.integers(-20, 30, n) * rng.choice([-1, 1], n)
    snap = rng.uniform(-60, 60, n)                       # snapshot-sized values
    small = rng.uniform(-1e-3, 1e-3, n)                  # pi components, velocities near 0
    ties = np.concatenate([
        (2 * rng.integers(0, 2**20, 20_000) + 1) / 2.0**9,   # exact ties at %.8f
        (2 * rng.integers(0, 2**12, 20_000) + 1) / 2.0**4,   # exact ties at %.3f
        (2 * rng.integers(50_000, 500_000, 20_000) + 1) / 2.0,  # exact ties at 6 significant digits
    ])
    near = np.concatenate([np.nextafter(ties, np.inf), np.nextafter(ties, -np.inf)])
    special = np.array([0.0, -0.0, np.inf, -np.inf, 1e-5, 1e-4, 9.999995e-5, 9.9999949e-5, 0.0001, 0.00001,
                        999999.5, 999999.4999999, 1e6, 1e15, 1e16, 1e22, 1e26, 1e27, 1e300, 5e-324, 2.2e-308,
                        123456.5, 0.5, 1.5, 2.5, -0.0005, -0.00000000499, 8.999999e6, 9.0e6, 9.1e6, 1.0 / 3,
                        2.0 / 3, 0.1, 0.2, 0.3, 1e-16, 1.0000005e-16, 9.99999e26])
    v = np.concatenate([wide, -wide[:1000], snap, small, ties, -ties, near, special])
    dec = rng.integers(-4, 6, 50_000)                    # decade edges: 10^k and its neighbours
    edge = 10.0 ** dec
    return np.concatenate([v, edge, np.nextafter(edge, 0), np.nextafter(edge, np.inf), edge * 0.9999995,
                           edge * 0.99999949999])


def test_formatter_matches_printf(harness):
    exe, d = harness
    v = _values()
    path = d / "v.bin"
    v.astype(np.float64).tofile(path)
    out = subprocess.run([str(exe), str(path)], check=True, capture_output=True, text=True).stdout.splitlines()
    assert len(out) == len(v)
    bad = []
    for x, line in zip(v.tolist(), out):
        want = "%g %.3f %.8f" % (x, x, x)
        if line != want:
            bad.append((repr(x), line, want))
            if len(bad) > 10:
                break
    assert not bad, bad
