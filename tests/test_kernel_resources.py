"""Register budgets of the fused step kernels at the benchmark shapes.

The kernels' speed rests on how many waves a CU holds, which the VGPR count
sets, and on not spilling to scratch.  Both are properties of the compiled
code, so the device code of the headline instantiations (256^3 x 17 and
512^3 x 37, FP32-mixed) is compiled here and the counts the compiler reports in
the code-object metadata are checked:
  * k_xline at 256 points: no spills, at most 96 VGPRs (5 waves per SIMD);
  * k_xline at 512 points: no spills;
  * k_ypass at 256 and 512 points: no spills;
  * k_zline at 256 points, with and without the Parseval sums: no spills, at
    most 168 VGPRs (3 waves per SIMD);
  * k_zline at 512 points with the 32-thread plan: at most 128 VGPRs, so two
    512-thread blocks share a CU (it spills a few registers by design).
Needs hipcc, not a GPU."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CSRC = os.path.join(ROOT, "phase-field-microstructure-evolution_amd", "csrc")

XLINE = ("template __global__ void pf::k_xline<float, {n}, 1, false, false>(pf::Geo, pf::UPar, pf::c2<float>*, "
         "pf::c2<float>*, float*, const pf::c2<float>*, double*, long, int, int, const pf::Member<float>*);")
YPASS = ("template __global__ void pf::k_ypass<float, {n}, false, false>(pf::c2<float>*, pf::c2<float>*, long, long, "
         "long, const pf::c2<float>*, const pf::Member<float>*, long);")
ZLINE = ("template __global__ void pf::k_zline<float, {n}, 0, true, {p}, false, {v}>(pf::Geo, pf::UPar, pf::KTab, "
         "pf::Bank<float>, pf::Bank<float>, pf::Bank<float>, pf::Bank<float>, double*, double*, "
         "const pf::c2<float>*, int, int, long, long, long, const pf::Member<float>*);")


@pytest.fixture(scope="module")
def resources(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("res")
    lines = ['#include "%s"' % os.path.join(CSRC, "pf_fused.hpp")]
    lines += [XLINE.format(n=n) for n in (256, 512)]
    lines += [YPASS.format(n=n) for n in (256, 512)]
    lines += [ZLINE.format(n=256, p=p, v=0) for p in ("false", "true")]
    lines += [ZLINE.format(n=512, p=p, v=1) for p in ("false", "true")]
    src = d / "res.hip"
    src.write_text("\n".join(lines) + "\n")
    asm = d / "res.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-S",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", str(src), "-o", str(asm)],
                   check=True, capture_output=True, timeout=900)
    out = {}
    # amdhsa.kernels metadata: one block per kernel with .name, .vgpr_count, .vgpr_spill_count
    for blk in re.split(r"\n\s+- \.", asm.read_text().split("amdhsa.kernels:", 1)[1]):
        name = re.search(r"\.name:\s+(\S+)", blk)
        vg = re.search(r"\.vgpr_count:\s+(\d+)", blk)
        sp = re.search(r"\.vgpr_spill_count:\s+(\d+)", blk)
        if name and vg and sp:
            out[name.group(1)] = (int(vg.group(1)), int(sp.group(1)))
    return out


def _find(res, kernel, n, *tail):
    # mangled names: _ZN2pf7k_xlineIfLi256ELi1ELb0ELb0EE..., template arguments in order
    pat = "_ZN2pf%d%sIfLi%dE%s" % (len(kernel), kernel, n, "".join(tail))
    hits = [v for k, v in res.items() if k.startswith(pat)]
    assert len(hits) == 1, (pat, sorted(res))
    return hits[0]


def test_xline_registers(resources):
    vg, sp = _find(resources, "k_xline", 256, "Li1ELb0ELb0E")
    assert sp == 0 and vg <= 96, (vg, sp)
    vg, sp = _find(resources, "k_xline", 512, "Li1ELb0ELb0E")
    assert sp == 0, (vg, sp)


def test_ypass_registers(resources):
    for n in (256, 512):
        vg, sp = _find(resources, "k_ypass", n, "Lb0ELb0E")
        assert sp == 0, (n, vg, sp)


def test_zline_registers(resources):
    for parseval in ("Lb0E", "Lb1E"):
        vg, sp = _find(resources, "k_zline", 256, "Li0ELb1E", parseval, "Lb0ELi0E")
        assert sp == 0 and vg <= 168, (parseval, vg, sp)
        vg, sp = _find(resources, "k_zline", 512, "Li0ELb1E", parseval, "Lb0ELi1E")
        assert vg <= 128, (parseval, vg, sp)
