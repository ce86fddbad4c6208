"""The buffer-descriptor loads and stores of the z-line kernel (csrc/pf_fft.hpp:
buf_ld, buf_st) must move a whole complex value: 8 bytes in FP32, 16 in FP64.
This compiler once narrowed them silently (a bit cast of a vector element read
element 0, after which the unused half of the load was dropped), so the device
code of a small kernel over the helpers is compiled here and its memory
instructions are checked by width.  Needs hipcc, not a GPU."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r'''
#include "%s"
__global__ void k64(double2* p, int n) {
  pf::BufRes r = pf::buf_res(p, n * 16);
  double2 w = pf::buf_ld<double>(r, threadIdx.x * 16, 4096);
  w.x *= 2.0; w.y *= 3.0;
  pf::buf_st<double>(r, threadIdx.x * 16, 8192, w);
}
__global__ void k32(float2* p, int n) {
  pf::BufRes r = pf::buf_res(p, n * 8);
  float2 w = pf::buf_ld<float, 0>(r, threadIdx.x * 8, 4096);
  w.x *= 2.0f; w.y *= 3.0f;
  pf::buf_st<float, 0>(r, threadIdx.x * 8, 8192, w);
}
'''


def test_buffer_accesses_move_whole_complex_values(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    src = tmp_path / "buf.hip"
    src.write_text(SRC % os.path.join(ROOT, "phase-field-microstructure-evolution_amd", "csrc", "pf_fft.hpp"))
    asm = tmp_path / "buf.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-S",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", str(src), "-o", str(asm)],
                   check=True, capture_output=True, timeout=300)
    text = asm.read_text()
    bodies = {}
    for name in ("k64", "k32"):
        start = re.search(r"^_Z\d*%s\w*:" % name, text, re.M).start()
        bodies[name] = text[start:text.index("s_endpgm", start)]
    ops = {k: re.findall(r"buffer_(load|store)_(dword\w*)", v) for k, v in bodies.items()}
    assert ops["k64"] == [("load", "dwordx4"), ("store", "dwordx4")], ops["k64"]
    assert ops["k32"] == [("load", "dwordx2"), ("store", "dwordx2")], ops["k32"]
    # the non-temporal hint is on the default form only
    assert " nt" in bodies["k64"] and " nt" not in bodies["k32"]
