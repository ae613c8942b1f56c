"""Host checks of the device arithmetic that has to round like the host's:

- the regulator's tanh (vh::glibc_tanh, cpu-vh_amd/csrc/vh_libm.hpp) against
  the host C library's tanh, which the reference calls
  (rhic/src/FullyDiscreteKurganovTadmorScheme.cpp:775,780), on 10^8 random
  arguments plus every double within 2^16 ulps of the branch thresholds:
  bit for bit;
- the reciprocal division with Markstein's correction (vh::DivFast,
  cpu-vh_amd/csrc/vh_device.hpp) against the division operator on 10^8 pairs:
  bit for bit;
- the fdlibm-method exp of the bulk-viscosity extension (vh::fdlibm_exp)
  against the host exp on 10^7 arguments: never more than one ulp apart
  (the two are different algorithms; the extension only needs the device and
  the oracle to share one).

tests/libm_check.cpp is compiled from the library's headers for the host
(hipcc, no device code, no contraction, like the device build).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("libm") / "libm_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off", "-mfma",
                    "-o", exe, os.path.join(HERE, "libm_check.cpp")], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("what", ["tanh", "div", "exp"])
def test_bitwise_against_host(checker, what):
    n = "10000000" if what == "exp" else "100000000"
    r = subprocess.run([checker, what, n], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("mismatches 0")


def test_checker_sees_a_wrong_tanh(checker, tmp_path):
    """The comparison fails on a planted error: the same program built with
    expm1's polynomial in fdlibm's original nested (Horner) order instead of
    the library's split order reports mismatches."""
    src = open(os.path.join(HERE, "..", "cpu-vh_amd", "csrc", "vh_libm.hpp")).read()
    bad = src.replace("const double r1 = R1 + h2 * R2 + h4 * R3;",
                      "const double r1 = one + hxs * (Q1 + hxs * (Q2 + hxs * (Q3 + hxs * (Q4 + hxs * Q5))));")
    assert bad != src
    d = tmp_path / "csrc"
    d.mkdir()
    (d / "vh_libm.hpp").write_text(bad)
    dev = open(os.path.join(HERE, "..", "cpu-vh_amd", "csrc", "vh_device.hpp")).read()
    (d / "vh_device.hpp").write_text(dev)
    t = tmp_path / "tests"
    t.mkdir()
    (t / "libm_check.cpp").write_text(open(os.path.join(HERE, "libm_check.cpp")).read()
                                       .replace("../cpu-vh_amd/csrc/", "../csrc/"))
    exe = str(tmp_path / "bad_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off", "-mfma",
                    "-o", exe, str(t / "libm_check.cpp")], check=True, capture_output=True)
    r = subprocess.run([exe, "tanh", "1000000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "mismatches 0" not in r.stdout


def test_checker_sees_a_wrong_exp(tmp_path):
    """A plausible slip in the exp (one polynomial coefficient's sign) is
    caught: the same checker built with it reports mismatches."""
    src = open(os.path.join(HERE, "..", "cpu-vh_amd", "csrc", "vh_libm.hpp")).read()
    bad = src.replace("P4 = -1.65339022054652515390e-06", "P4 = 1.65339022054652515390e-06")
    assert bad != src
    d = tmp_path / "csrc"
    d.mkdir()
    (d / "vh_libm.hpp").write_text(bad)
    (d / "vh_device.hpp").write_text(open(os.path.join(HERE, "..", "cpu-vh_amd", "csrc", "vh_device.hpp")).read())
    t = tmp_path / "tests"
    t.mkdir()
    (t / "libm_check.cpp").write_text(open(os.path.join(HERE, "libm_check.cpp")).read()
                                       .replace("../cpu-vh_amd/csrc/", "../csrc/"))
    exe = str(tmp_path / "bad_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off", "-mfma",
                    "-o", exe, str(t / "libm_check.cpp")], check=True, capture_output=True)
    r = subprocess.run([exe, "exp", "1000000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "mismatches 0" not in r.stdout
