"""End to end through the Fortran host driver var_diff_hip.x: initial
condition (circle, or a restart file), precipitate insertion at ppt_int, GPU
timesteps, summary.dat rows and data_save checkpoints, against the oracle
driven through the same sequence (oracle/ics.py for the host parts, the oracle
timestep for the rest).

summary.dat is written with 8 significant digits (format e15.7), so rows are
compared at 1e-6 relative (FP64 forms) / 1e-4 (FP32-mixed); checkpoints at
1e-11 / 1e-4 absolute."""
import os

import numpy as np
import pytest

import oracle
from oracle import ics
from driver_util import DRIVER, cfg_of, run_driver, write_input

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRIVER), reason="var_diff_hip.x not built")]


def _circle_state(p):
    phi = np.zeros((p["Nz"], p["Ny"], p["Nx"], p["var"]))
    con = np.full((p["Nz"], p["Ny"], p["Nx"]), p["c_mat"])
    rng = oracle.Ran2(p["iseed"])
    ics.circle(cfg_of(p), phi, con, rng)
    return phi, con, rng


def _oracle_run(p, phi, con, rng):
    cfg = cfg_of(p)
    prm = oracle.make_params(t_step=p["t_step"], grad_coeff_phi=p["grad_coeff_phi"],
                             grad_coeff_c=p["grad_coeff_c"], D_S=p["D_S"], D_L=p["D_L"],
                             D_gb=p["d_gb"], mob_phi=p["mob_phi"], sigma_1=p["sigma_1"],
                             sigma_2=p["sigma_2"], am=p["am"], ap=p["ap"],
                             con_0_mat=p["con_0_mat"], con_0_ppt=p["con_0_ppt"])
    rows, snaps = {}, {}
    for step in range(1, p["N_step"] + 1):
        ext = None
        if step == p["ppt_int"]:
            # the row of step ppt_int carries the extrema the reference computed
            # at the end of the previous step, before the insertion
            # (src/var_diff.f90:479, 1071, 1076-1094)
            if step % 2 == 0:
                ext = oracle.extrema(phi, con)
            ics.insert(cfg, phi, rng, max_draws=100_000)
        d = oracle.run(prm, phi, con, 1)[0]
        if ext is not None:
            d.update(ext)
        if step % 2 == 0:
            rows[step] = [d[k] for k in oracle.DIAG_FIELDS]
        if step % p["ifreq"] == 0:
            snaps[step] = (phi.copy(), con.copy())
    return rows, snaps


def _check(out, p, mode, tol, rows, snaps):
    got = {}
    for line in open(out / "summary.dat"):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        got[int(f[0])] = [float(x) for x in f[1:]]
    assert sorted(got) == sorted(rows)
    rtol, atol = (1e-6, 1e-12) if mode != "fp32_mixed" else (1e-4, 1e-6)
    for s in rows:
        for a, b in zip(got[s], rows[s]):
            assert abs(a - b) <= rtol * abs(b) + atol, (s, got[s], rows[s])
    for s, (ophi, ocon) in snaps.items():
        path = out / f"step_{s:07d}" / "data_save.0001"
        _, _, phi, con = ics.read_checkpoint(str(path), p["Nx"], p["Ny"], p["Nz"], p["var"])
        assert np.abs(phi - ophi).max() < tol and np.abs(con - ocon).max() < tol, s
    assert sum(1 for l in open(out / "time_step.dat") if not l.startswith("#")) == p["N_step"]
    assert sum(1 for l in open(out / "timing_data.dat") if not l.startswith("#")) == p["N_step"]


MODES = [("fp64_faithful", 1e-11), ("fp64_reduced", 1e-11), ("fp32_mixed", 1e-4)]


@pytest.mark.parametrize("mode,tol", MODES)
def test_driver_circle(tmp_path, mode, tol):
    inp = tmp_path / "input.txt"
    p = write_input(inp, Nx=24, Ny=32, Nz=16, var=3, N_step=6, ifreq=3, ppt_int=0, num_ppt=2,
                    grn_rad=6, d_gb=0.2)
    out = tmp_path / "out"
    # one form with the per-stage timing columns (steps issued one at a time),
    # the others batched between checkpoints
    timing = mode == "fp64_reduced"
    env = {"PF_MODE": mode, **({"PF_TIMING": "1"} if timing else {})}
    print("driver circle", mode, "timing" if timing else "", flush=True)
    r = run_driver(str(tmp_path), str(inp), str(out), env=env, timeout=60)
    assert r.returncode == 0, r.stderr
    _check(out, p, mode, tol, *_oracle_run(p, *_circle_state(p)))
    cols = [l.split() for l in open(out / "timing_data.dat") if not l.startswith("#")]
    upd = [float(c[13]) for c in cols]  # loop 13: the spectral update
    assert all(u > 0 for u in upd) if timing else all(u == 0 for u in upd)


def restart_state(pf):
    """Three-grain polycrystal folded into two grain fields, precipitate field
    exactly zero (as in the reference's polycrystal runs before insertion)."""
    g, _ = pf.ic_voronoi(24, 24, 24, 3, 6, -11)
    phi0 = np.concatenate([g[..., :2], np.zeros_like(g[..., :1])], axis=-1)
    phi0[..., 0] += g[..., 2]
    return phi0, np.full((24, 24, 24), 0.4)


RESTART = dict(nrun=2, Nx=24, Ny=24, Nz=24, var=3, N_step=8, ifreq=4, ppt_int=5, num_ppt=3,
               ppt_rad=(2, 2, 2), d_gb=0.3, iseed=-5)


@pytest.mark.parametrize("ppt_int", [5, 6])
@pytest.mark.parametrize("mode,tol", MODES)
def test_driver_restart_with_insertion(tmp_path, pf, mode, tol, ppt_int):
    """Restart (nrun = 2), then insertion at ppt_int once the grain boundaries
    have become diffuse (the insertion only accepts centres with
    sum phi_grain^2 <= 0.8 and no precipitate in reach).  With an even ppt_int
    the row of that step is reported and must carry the pre-insertion extrema."""
    phi0, con0 = restart_state(pf)
    ics.write_restart(str(tmp_path / "data_rerun.0001"), phi0, con0)
    inp = tmp_path / "input.txt"
    p = write_input(inp, **dict(RESTART, ppt_int=ppt_int))
    out = tmp_path / "out"
    print("driver restart+insert", mode, "ppt_int", ppt_int, flush=True)
    r = run_driver(str(tmp_path), str(inp), str(out), env={"PF_MODE": mode}, timeout=60)
    assert r.returncode == 0, r.stderr
    rows, snaps = _oracle_run(p, phi0.copy(), con0.copy(), oracle.Ran2(p["iseed"]))
    _check(out, p, mode, tol, rows, snaps)
    assert (snaps[8][0][..., 2] > 0.5).sum() > 0  # precipitates were inserted


def test_driver_error_log_against_fp64(tmp_path, pf):
    """PF_ERROR_LOG with the FP32-mixed form: a second, fp64_faithful run of the
    same input alongside, and per checkpoint the per-field max and RMS
    difference in error_vs_fp64.dat.  The logged values must be the differences
    of the FP32 checkpoint from the oracle (which the FP64 run matches to 1e-11)."""
    phi0, con0 = restart_state(pf)
    ics.write_restart(str(tmp_path / "data_rerun.0001"), phi0, con0)
    inp = tmp_path / "input.txt"
    p = write_input(inp, **dict(RESTART, ppt_int=6))
    out = tmp_path / "out"
    print("driver error log fp32_mixed", flush=True)
    r = run_driver(str(tmp_path), str(inp), str(out), env={"PF_MODE": "fp32_mixed", "PF_ERROR_LOG": "1"},
                   timeout=60)
    assert r.returncode == 0, r.stderr
    rows, snaps = _oracle_run(p, phi0.copy(), con0.copy(), oracle.Ran2(p["iseed"]))
    _check(out, p, "fp32_mixed", 1e-4, rows, snaps)
    lines = [l.split() for l in open(out / "error_vs_fp64.dat") if not l.startswith("#")]
    assert [int(l[0]) for l in lines] == sorted(snaps)
    for l in lines:
        s = int(l[0])
        vals = np.array([float(x) for x in l[1:]]).reshape(p["var"] + 1, 2)
        _, _, phi, con = ics.read_checkpoint(str(out / f"step_{s:07d}" / "data_save.0001"),
                                             p["Nx"], p["Ny"], p["Nz"], p["var"])
        ophi, ocon = snaps[s]
        diffs = [con - ocon] + [phi[..., i] - ophi[..., i] for i in range(p["var"])]
        for f, dd in enumerate(diffs):
            assert vals[f, 0] == pytest.approx(np.abs(dd).max(), rel=1e-4, abs=1e-11), (s, f)
            assert vals[f, 1] == pytest.approx(np.sqrt((dd * dd).mean()), rel=1e-4, abs=1e-11), (s, f)
        assert 0 < vals[:, 0].max() < 1e-5
