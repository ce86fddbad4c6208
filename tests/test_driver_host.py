"""CPU tests of the Fortran host driver (var_diff_hip.x): input parsing,
gamma_matrix.dat, the initial conditions and the restart reader, run with
PF_DRY_RUN so that no GPU is touched.  The checker is the restatement in
oracle/ics.py of src/var_diff.f90:229-461 with the oracle's ran_2."""
import os

import numpy as np
import pytest

import oracle
from oracle import ics
from driver_util import DRIVER, cfg_of, run_driver, write_input

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="var_diff_hip.x not built")


def _dry(tmp_path, check_state=True, **kw):
    inp = tmp_path / "input.txt"
    p = write_input(inp, **kw)
    out = tmp_path / "out"
    r = run_driver(str(tmp_path), str(inp), str(out), env={"PF_DRY_RUN": "1"})
    assert r.returncode == 0, r.stderr
    if not check_state:
        return p, None, None, out, r
    ist, ien, phi, con = ics.read_checkpoint(str(out / "step_0000000" / "data_save.0001"),
                                              p["Nx"], p["Ny"], p["Nz"], p["var"])
    assert list(ist) == [1, 1, 1] and list(ien) == [p["Nx"], p["Ny"], p["Nz"]]
    return p, phi, con, out, r


@pytest.mark.parametrize("shape,num_ppt", [((24, 24, 24), 3), ((32, 16, 24), 2)])
def test_circle_ic_matches_restatement(tmp_path, shape, num_ppt):
    p, phi, con, _, r = _dry(tmp_path, Nx=shape[0], Ny=shape[1], Nz=shape[2], num_ppt=num_ppt,
                             grn_rad=5, start="circle")
    ephi = np.zeros_like(phi)
    econ = np.full_like(con, p["c_mat"])
    ics.circle(cfg_of(p), ephi, econ, oracle.Ran2(p["iseed"]))
    assert np.array_equal(phi, ephi)
    assert np.array_equal(con, econ)
    assert (phi[..., 2] == 1.0).sum() > 0 and r.stdout.count("nppt=") == num_ppt


def test_plane_ic_matches_restatement(tmp_path):
    p, phi, con, _, _ = _dry(tmp_path, Nx=24, Ny=32, Nz=16, start="plane", i_ppt=10, j_ppt=20,
                             k_ppt=8, ppt_rad=(3, 2, 2))
    ephi = np.zeros_like(phi)
    ics.plane(cfg_of(p), ephi)
    assert np.array_equal(phi, ephi)
    assert np.all(con == p["c_mat"])


def test_restart_reader_roundtrip(tmp_path):
    import __graft_entry__
    pf = __graft_entry__.load_package()
    phi0, con0 = pf.ic_voronoi(16, 24, 12, 4, 5)
    ics.write_restart(str(tmp_path / "data_rerun.0001"), phi0, con0)
    p, phi, con, _, _ = _dry(tmp_path, nrun=2, Nx=16, Ny=24, Nz=12, var=4)
    assert np.array_equal(phi, phi0) and np.array_equal(con, con0)


def test_gamma_matrix_file(tmp_path):
    p, _, _, out, _ = _dry(tmp_path, var=5, sigma_1=1.25, sigma_2=0.75)
    rows = [l.split() for l in open(out / "gamma_matrix.dat") if l.strip()]
    got = np.array([[float(x) for x in r[1:]] for r in rows])
    assert [int(r[0]) for r in rows] == list(range(1, 6))
    assert np.array_equal(got, ics.gamma(5, 1.25, 0.75))


def test_input_comment_and_headers(tmp_path):
    """The '/' ending each record makes the trailing text a comment; output
    files carry the reference's header lines."""
    _, _, _, out, _ = _dry(tmp_path)
    assert open(out / "summary.dat").readline().startswith("# step number, f_int_tot")
    assert open(out / "time_step.dat").readline().startswith("# step number")
    assert open(out / "timing_data.dat").readline().startswith("# step number, loop 1 avg")


def test_checkpoint_reader_rejects_truncation(tmp_path):
    ics.write_fortran_records(str(tmp_path / "f"), [b"\x00" * 24, b"\x01" * 80])
    data = open(tmp_path / "f", "rb").read()
    open(tmp_path / "g", "wb").write(data[:-6])
    with pytest.raises(Exception):
        ics.read_fortran_records(str(tmp_path / "g"))


def test_record_over_2gib_uses_subrecords(tmp_path):
    """At 256^3 x 17 fields the phi+con record is 2.28 GB: it must be split into
    subrecords (a single 4-byte marker cannot hold its length).  Restart from a
    file written by the restatement, write it back, and compare byte for byte."""
    n, var = 256, 16
    rng = np.random.default_rng(3)
    phi0 = rng.random((n, n, n, var))
    con0 = rng.random((n, n, n))
    src = tmp_path / "data_rerun.0001"
    ics.write_restart(str(src), phi0, con0)
    del phi0, con0
    _, _, _, out, _ = _dry(tmp_path, nrun=2, Nx=n, Ny=n, Nz=n, var=var, check_state=False)
    dst = out / "step_0000000" / "data_save.0001"
    assert os.path.getsize(dst) == os.path.getsize(src) == 24 + 8 + 8 * 2 + 8 * n ** 3 * (var + 1)
    with open(src, "rb") as a, open(dst, "rb") as b:
        while True:
            x, y = a.read(1 << 26), b.read(1 << 26)
            assert x == y
            if not x:
                break
    with open(dst, "rb") as f:
        f.seek(32)
        import struct
        assert struct.unpack("<i", f.read(4))[0] == -ics.MAX_SUBRECORD


def test_truncated_restart_is_rejected(tmp_path):
    phi0 = np.zeros((8, 8, 8, 3))
    con0 = np.full((8, 8, 8), 0.3)
    src = tmp_path / "data_rerun.0001"
    ics.write_restart(str(src), phi0, con0)
    data = open(src, "rb").read()
    open(src, "wb").write(data[:-100])
    inp = tmp_path / "input.txt"
    write_input(inp, nrun=2, Nx=8, Ny=8, Nz=8, var=3)
    r = run_driver(str(tmp_path), str(inp), str(tmp_path / "out"), env={"PF_DRY_RUN": "1"})
    assert r.returncode != 0 and "truncated" in r.stderr


def test_polycryst_ic_from_grain_map(tmp_path):
    """start = 'polycryst': one 'i j k ivar' line per point, any order
    (src/var_diff.f90:232-249)."""
    nx, ny, nz, var = 12, 8, 10, 5
    rng = np.random.default_rng(4)
    ivar = rng.integers(1, var + 1, size=(nz, ny, nx))
    pts = [(i + 1, j + 1, k + 1, int(ivar[k, j, i])) for k in range(nz) for j in range(ny) for i in range(nx)]
    order = rng.permutation(len(pts))
    gm = tmp_path / "grains.txt"
    gm.write_text("".join("%d %d %d %d\n" % pts[o] for o in order))
    inp = tmp_path / "input.txt"
    p = write_input(inp, Nx=nx, Ny=ny, Nz=nz, var=var, start="polycryst")
    out = tmp_path / "out"
    r = run_driver(str(tmp_path), str(inp), str(out), env={"PF_DRY_RUN": "1", "PF_POLYCRYST_FILE": str(gm)})
    assert r.returncode == 0, r.stderr
    _, _, phi, con = ics.read_checkpoint(str(out / "step_0000000" / "data_save.0001"), nx, ny, nz, var)
    expect = np.zeros((nz, ny, nx, var))
    for (i, j, k, v) in pts:
        expect[k - 1, j - 1, i - 1, v - 1] = 1.0
    assert np.array_equal(phi, expect) and np.all(con == p["c_mat"])
    # a grain index beyond var is an error, not a stray write
    gm.write_text("1 1 1 %d\n" % (var + 1) + "".join("%d %d %d %d\n" % q for q in pts[1:]))
    r = run_driver(str(tmp_path), str(inp), str(out), env={"PF_DRY_RUN": "1", "PF_POLYCRYST_FILE": str(gm)})
    assert r.returncode != 0 and "outside" in r.stderr


def test_checkpoint_read_by_fortran_runtime(tmp_path):
    """The checkpoint read back with plain Fortran sequential reads (the
    post-processor's way), compiled here with amdflang."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or "/opt/rocm/llvm/bin/amdflang"
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    exe = tmp_path / "read_checkpoint"
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "read_checkpoint.f90")
    subprocess.run([fc, "-O1", src, "-o", str(exe)], check=True, capture_output=True, cwd=str(tmp_path))
    p, phi, con, out, _ = _dry(tmp_path, Nx=24, Ny=16, Nz=8, var=4, grn_rad=4, num_ppt=2, start="circle")
    r = subprocess.run([str(exe), str(out / "step_0000000" / "data_save.0001"), "24", "16", "8", "4"],
                       check=True, capture_output=True, text=True)
    lines = r.stdout.split("\n")
    assert [int(x) for x in lines[0].split()] == [1, 1, 1, 24, 16, 8]
    s = [float(x) for x in lines[1].split()]
    assert s[0] == pytest.approx(phi.sum(), rel=1e-13) and s[1] == pytest.approx(con.sum(), rel=1e-13)
    assert s[2] == pytest.approx((phi[..., 3] * con).sum(), rel=1e-13)
