"""CPU check of the register-resident Stockham schedule in csrc/pf_fft2.hpp.

tests/cpu/fft2_emul.cpp compiles the kernel header for the host and replays,
for every supported transform length, each thread's butterflies with the LDS
exchanges in between; the result is compared with a long-double DFT.  This
pins the plan tables (slot <-> element maps, twiddle indices) independently of
the GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def emul_output(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("emul") / "fft2_emul")
    cmd = [HIPCC, "-O1", "-std=c++17", "--offload-host-only", "--offload-arch=gfx950",
           os.path.join(ROOT, "tests", "cpu", "fft2_emul.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=300).stdout
    return {(int(l.split()[0]), int(l.split()[1])): l.split() for l in out.strip().splitlines()}


LENGTHS = [2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 512]
# alternative radix sequences (plan variants 1, 2), candidates of the tuner
VARIANTS = [(n, v) for v in (1, 2) for n in (64, 128, 160, 192, 256, 512)]


@pytest.mark.parametrize("n,v", [(n, 0) for n in LENGTHS] + VARIANTS)
def test_schedule_matches_dft(emul_output, n, v):
    _, _, npass, ef, ei, inplace, _t = emul_output[(n, v)]
    assert int(npass) <= 3
    # double-precision butterflies: error grows ~ log n * eps * |x|_1
    assert float(ef) < 1e-13 * n and float(ei) < 1e-13 * n
    # the last pass writes each thread's outputs to its own input elements
    # (k_xline writes the transform back without a barrier on this)
    assert inplace == "1"


@pytest.mark.parametrize("n", [4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 256])
def test_swizzled_exchange_matches_dft(emul_output, n):
    """The x-line rows store element e at lds_swz(e) = e + e/16; the transform
    with its exchanges through that layout still computes the DFT, and lds_swz
    maps [0, n) one-to-one into the n + n/16 elements of a row."""
    _, _, npass, ef, ei, inplace, perm = emul_output[(n, 9)]
    assert float(ef) < 1e-13 * n and float(ei) < 1e-13 * n
    assert perm == "1"


@pytest.mark.parametrize("n", [64, 128, 256])
def test_wave_wide_plans_match_dft(emul_output, n):
    """Plan variant 3: the x line's lone job in a block's last wave, transformed
    by up to 64 threads (radix 2 at 64 and 128 points, radix 4 at 256)."""
    _, _, npass, ef, ei, inplace, nthreads = emul_output[(n, 3)]
    assert float(ef) < 1e-13 * n and float(ei) < 1e-13 * n
    assert inplace == "1" and int(nthreads) <= 64


def test_variants_differ_from_the_default_plan(emul_output):
    for n, v in VARIANTS:
        assert emul_output[(n, v)][6] != emul_output[(n, 0)][6] or \
            emul_output[(n, v)][2] != emul_output[(n, 0)][2], (n, v)
