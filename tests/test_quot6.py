"""The face kernels' quotient by 6 (fvg::quot6, three operations in place of the ten of the IEEE division) against the
IEEE division itself, on the host, from the same header source. WENO5-Z's candidate values divide by 6
(Schemes.cpp:343-345), so the FP64 face kernels give the reference's bits only if quot6 equals x / 6.0 everywhere it
is used: tests/quot6_check.cpp tries random significands over its range, exact multiples of 6 and their neighbours,
quotients next to powers of two, zeros and the range ends, and WENO5-Z numerators of stencil values that
range_fits_quot6 (the k_cells check that selects it) accepts."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("quot6") / "quot6_check")
    subprocess.run([HIPCC, "-x", "hip", "-O2", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    os.path.join(HERE, "quot6_check.cpp"), "-o", exe], check=True, capture_output=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_quot6_equals_the_ieee_division(checker, seed):
    r = subprocess.run([checker, "3000000", str(seed)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "total differing: 0" in r.stdout
    assert "numerators outside quot6's range: 0" in r.stdout
    assert "range_fits_quot6 bounds: ok" in r.stdout
    # below the range the three-operation form is not the division: the reason k_cells checks the range
    assert "x / 6 = 0x0.0000000000002p-1022, quot6 = 0x0.0000000000001p-1022" in r.stdout
