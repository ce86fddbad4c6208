"""GPU timestep parity against the CPU FP64 oracle (restatement of
src/var_diff.f90:474-1118) on seeded inputs small enough for the oracle.

Tolerances (about 30x / 4x the largest difference measured over these shapes:
FP64 3.7e-15, FP32 4.7e-7 per field): FP64 forms 1e-13 relative per field
(max-norm), on the summary energies and on the extrema; FP32-mixed 2e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VOR = dict(t_step=0.005, grad_coeff_phi=1.0, grad_coeff_c=1.0)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def energy_scale(o, k):
    """The magnitude a summary energy's error is relative to: its own for the
    three sums, and for e_tot = f_int + f_ch + f_eta the sum of their magnitudes
    (the terms can nearly cancel: e_tot = -16 from terms of ~1e4 at 512 x 16 x 16
    x 37, where each term's rounding is what e_tot inherits)."""
    if k == "e_tot":
        return max(abs(o["f_int"]) + abs(o["f_ch"]) + abs(o["f_eta"]), 1.0)
    return max(abs(o[k]), 1.0)


def _compare(pf, oracle, shape, var, nseeds, nsteps, mode, params, tol_state, tol_e, tuning=None):
    nx, ny, nz = shape
    phi, con = pf.ic_voronoi(nx, ny, nz, var, nseeds)
    with pf.Simulation(pf.make_params(**params), nx, ny, nz, var, mode) as sim:
        if tuning:
            sim.set_tuning(**tuning)
        sim.put_state(phi, con)
        dg = sim.step(nsteps, diag_every=2)
        gphi, gcon = sim.get_state()
    ophi, ocon = phi.copy(), con.copy()
    od = oracle.run(oracle.make_params(**params), ophi, ocon, nsteps)
    e_worst = max((abs(d[k] - od[d["step"] - 1][k]) / energy_scale(od[d["step"] - 1], k)
                   for d in dg for k in ("f_int", "f_ch", "f_eta", "e_tot")), default=0.0)
    x_worst = max((abs(d[k] - od[d["step"] - 1][k]) for d in dg
                   for k in ("c_max", "c_min", "phi_min", "phi_3_max")), default=0.0)
    print(f"{mode} {shape} var={var} steps={nsteps}: phi {_rel(gphi, ophi):.2e} con {_rel(gcon, ocon):.2e} "
          f"energies {e_worst:.2e} extrema {x_worst:.2e}")
    assert np.isfinite(gphi).all() and np.isfinite(gcon).all()
    assert _rel(gphi, ophi) < tol_state
    assert _rel(gcon, ocon) < tol_state
    assert len(dg) == nsteps // 2
    for d in dg:
        o = od[d["step"] - 1]
        for k in ("f_int", "f_ch", "f_eta", "e_tot"):
            assert abs(d[k] - o[k]) <= tol_e * energy_scale(o, k), (k, d[k], o[k])
        for k in ("c_max", "c_min", "phi_min", "phi_3_max"):
            assert abs(d[k] - o[k]) <= tol_state, (k, d[k], o[k])


@pytest.mark.parametrize("shape,var,nseeds", [((16, 16, 16), 3, 6), ((24, 16, 32), 4, 8),
                                              ((32, 32, 32), 16, 16)])
@pytest.mark.parametrize("dgb", [0.0, 0.3])
def test_fp64_faithful_vs_oracle(pf, oracle, shape, var, nseeds, dgb):
    _compare(pf, oracle, shape, var, nseeds, 10, "fp64_faithful", dict(VOR, D_gb=dgb), 1e-13, 1e-13)


@pytest.mark.parametrize("shape,var,nseeds", [((16, 16, 16), 3, 6), ((24, 16, 32), 4, 8),
                                              ((32, 32, 32), 16, 16)])
@pytest.mark.parametrize("dgb", [0.0, 0.3])
def test_fp64_reduced_vs_oracle(pf, oracle, shape, var, nseeds, dgb):
    _compare(pf, oracle, shape, var, nseeds, 10, "fp64_reduced", dict(VOR, D_gb=dgb), 1e-13, 1e-13)


@pytest.mark.parametrize("dgb", [0.0, 0.3])
def test_fp32_mixed_vs_oracle(pf, oracle, dgb):
    _compare(pf, oracle, (32, 32, 32), 4, 8, 20, "fp32_mixed", dict(VOR, D_gb=dgb), 2e-6, 2e-6)


# One or two order parameters: the h sum runs over no field (var = 1) or one
# (var = 2), the gamma matrix has no sigma_1 pair, and the x-line blocks hold 2
# or 3 fields.
@pytest.mark.parametrize("var", [1, 2])
@pytest.mark.parametrize("mode,tol_s,tol_e", [("fp64_faithful", 1e-13, 1e-13), ("fp64_reduced", 1e-13, 1e-13),
                                              ("fp32_mixed", 2e-6, 2e-6)])
def test_few_order_parameters_vs_oracle(pf, oracle, var, mode, tol_s, tol_e):
    _compare(pf, oracle, (16, 24, 16), var, 4, 10, mode, dict(VOR, D_gb=0.3), tol_s, tol_e)


@pytest.mark.parametrize("mode", ["fp64_faithful", "fp64_reduced", "fp32_mixed"])
def test_diagnose_matches_next_step_record(pf, mode):
    """pf_diagnose reports the start-of-step diagnostics of the current state, i.e.
    the record pf_step would produce for the next step."""
    n, var = 16, 4
    phi, con = pf.ic_voronoi(n, n, n, var, 6)
    params = pf.make_params(t_step=0.005, grad_coeff_phi=1.0, grad_coeff_c=1.0)
    with pf.Simulation(params, n, n, n, var, mode) as sim:
        sim.put_state(phi, con)
        sim.step(3)
        d0 = sim.diagnose()
        d1 = sim.step(1, diag_every=1)[0]
    tol = 1e-12 if mode != "fp32_mixed" else 1e-5
    for k in ("f_int", "f_ch", "f_eta", "e_tot", "c_max", "c_min", "phi_min", "phi_3_max", "c_mean"):
        assert abs(d0[k] - d1[k]) <= tol * max(1.0, abs(d1[k])), (k, d0[k], d1[k])
    assert d0["step"] == d1["step"] == 4


@pytest.mark.parametrize("mode", ["fp64_faithful", "fp64_reduced"])
def test_split_run_equals_continuous_run(pf, mode):
    """2k steps == k steps, get_state, new handle, put_state, k steps (FP64 forms)."""
    n, var, k = 16, 3, 6
    phi, con = pf.ic_voronoi(n, n, n, var, 5)
    params = pf.make_params(t_step=0.005, grad_coeff_phi=1.0, grad_coeff_c=1.0, D_gb=0.2)
    with pf.Simulation(params, n, n, n, var, mode) as a:
        a.put_state(phi, con)
        a.step(2 * k)
        pa, ca = a.get_state()
    with pf.Simulation(params, n, n, n, var, mode) as b:
        b.put_state(phi, con)
        b.step(k)
        pm, cm = b.get_state()
    with pf.Simulation(params, n, n, n, var, mode) as c:
        c.put_state(pm, cm)
        c.step(k)
        pb, cb = c.get_state()
    if mode == "fp64_faithful":
        assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
    else:  # the carried spectrum is rebuilt from the real state at put_state
        assert np.abs(pa - pb).max() < 1e-13 and np.abs(ca - cb).max() < 1e-13


def test_fp32_long_x_lines_many_fields(pf, oracle):
    """x lines of 512 points with 37 fields: one 1024-thread block per line and
    80 KiB of LDS (the 512^3 x 37 configuration's x-line shape, on a 512 x 16 x 16
    grid the oracle can follow)."""
    _compare(pf, oracle, (512, 16, 16), 36, 40, 6, "fp32_mixed", dict(VOR, D_gb=0.0), 2e-6, 2e-6)


def test_fp32_long_x_lines_several_groups_per_block(pf, oracle):
    """As above with 1024 lines over 256 blocks: each block takes four line groups
    in turn (with D_gb != 0, the mobility chain's one-field blocks too)."""
    _compare(pf, oracle, (512, 32, 32), 36, 40, 4, "fp32_mixed", dict(VOR, D_gb=0.25), 2e-6, 2e-6,
             tuning=dict(xline_grid=256))


def test_fp64_long_x_lines_many_fields(pf, oracle):
    """FP64 x lines of 512 points with 37 fields: 640-thread blocks and unpadded
    rows (156 KiB of LDS), against the oracle."""
    _compare(pf, oracle, (512, 16, 16), 36, 40, 4, "fp64_reduced", dict(VOR, D_gb=0.25), 1e-13, 1e-13)


def test_fp64_too_many_fields_is_an_error(pf):
    """41 fields of 512-point FP64 lines need 656 threads, past the 640 an FP64
    block holds: refused when the handle is created."""
    with pytest.raises(pf.PfError) as e:
        with pf.Simulation(pf.make_params(**VOR), 512, 8, 8, 40, "fp64_reduced"):
            pass
    assert e.value.code == pf.PF_ENOTIMPL


@pytest.mark.parametrize("mode", ["fp64_faithful", "fp64_reduced"])
def test_fp64_500_steps_vs_oracle(pf, oracle, mode):
    """500 steps of a 16^3 x 5-field polycrystal with D_gb != 0: per-field max
    relative difference to the oracle <= 1e-12."""
    n, var = 16, 4
    prm = dict(VOR, D_gb=0.3)
    phi, con = pf.ic_voronoi(n, n, n, var, 6)
    with pf.Simulation(pf.make_params(**prm), n, n, n, var, mode) as sim:
        sim.put_state(phi, con)
        sim.step(500, diag_every=0)
        gphi, gcon = sim.get_state()
    ophi, ocon = phi.copy(), con.copy()
    oracle.run(oracle.make_params(**prm), ophi, ocon, 500)
    errs = [_rel(gphi[..., i], ophi[..., i]) for i in range(var)] + [_rel(gcon, ocon)]
    print(mode, "500-step per-field rel diff", ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-12


def test_fp32_error_after_1000_steps(pf):
    """FP32-mixed against FP64 over 1000 steps of a 32^3 x 5-field polycrystal,
    printed every 100 steps next to the difference between the two FP64 forms
    (which differ only in rounding, ~1e-16 per step): the ratio of the two
    curves is the rounding unit's ratio, so both measure how strongly these
    dynamics amplify rounding.  Bound 3e-5 per field (measured 1.3e-5 in c;
    the benchmark workload at 256^3 x 17 stays below 1e-5, test_gpu_configs).
    The FP32 state multiplier is applied as a float hi+lo pair, so its rounding
    does not bias every step the same way; without that the error grows to ~1e-4."""
    n, var = 32, 4
    phi, con = pf.ic_voronoi(n, n, n, var, 8)
    prm = pf.make_params(**VOR)
    sims = {m: pf.Simulation(prm, n, n, n, var, m) for m in ("fp32_mixed", "fp64_reduced", "fp64_faithful")}
    try:
        for s in sims.values():
            s.put_state(phi, con)
        for ck in range(1, 11):
            st = {}
            for m, s in sims.items():
                s.step(100)
                st[m] = s.get_state()
            (pa, ca), (pb, cb), (pc, cc) = st["fp32_mixed"], st["fp64_reduced"], st["fp64_faithful"]
            e32 = [float(np.abs(ca - cb).max())] + [float(np.abs(pa[..., i] - pb[..., i]).max()) for i in range(var)]
            e64 = [float(np.abs(cc - cb).max())] + [float(np.abs(pc[..., i] - pb[..., i]).max()) for i in range(var)]
            print(f"32^3 step {ck * 100}: |fp32 - fp64| {max(e32):.2e} (c {e32[0]:.2e}), "
                  f"|fp64 faithful - reduced| {max(e64):.2e}, ratio {max(e32) / max(max(e64), 1e-300):.2e}")
    finally:
        for s in sims.values():
            s.close()
    print("fp32 vs fp64 after 1000 steps, per-field max", ["%.1e" % e for e in e32])
    assert max(e32) <= 3e-5


def test_ensemble_step_equals_individual_steps(pf):
    """pf_ensemble_step interleaves the members' steps on their own streams; each
    member's result is bit-identical to stepping it alone."""
    shapes = [(16, 16, 16, 3, "fp32_mixed"), (24, 16, 32, 4, "fp64_reduced"), (32, 32, 32, 3, "fp64_faithful")]
    prm = pf.make_params(**dict(VOR, D_gb=0.2))
    ics_ = [pf.ic_voronoi(nx, ny, nz, var, 5, -(i + 1)) for i, (nx, ny, nz, var, _) in enumerate(shapes)]
    alone = []
    for (nx, ny, nz, var, mode), (phi, con) in zip(shapes, ics_):
        with pf.Simulation(prm, nx, ny, nz, var, mode) as s:
            s.put_state(phi, con)
            s.step(7)
            alone.append(s.get_state())
    sims = []
    try:
        for (nx, ny, nz, var, mode), (phi, con) in zip(shapes, ics_):
            s = pf.Simulation(prm, nx, ny, nz, var, mode)
            s.put_state(phi, con)
            sims.append(s)
        pf.ensemble_step(sims, 7)
        for s, (pa, ca) in zip(sims, alone):
            pb, cb = s.get_state()
            assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
    finally:
        for s in sims:
            s.close()
    with pf.Simulation(prm, 16, 16, 16, 3, "fp32_mixed") as s:
        with pytest.raises(pf.PfError):
            pf.ensemble_step([s, s], 1)  # a handle may appear once


def test_batched_ensemble_equals_individual_steps(pf):
    """Members of one grid shape in the FP32-mixed form with constant mobility
    step together, one launch per stage over all of them, each with its own
    parameters (kappa, M, L as a parameter sweep draws them); every member's
    result is bit-identical to stepping it alone.  Two shapes form two groups;
    a member with D_gb != 0 and a faithful member step on their own."""
    cases = [(16, 16, 16, 3, 0.5, 0.8, 1.3, 0.0), (16, 16, 16, 3, 1.7, 1.2, 0.6, 0.0),
             (16, 16, 16, 3, 0.9, 0.5, 1.9, 0.0), (24, 16, 32, 4, 1.1, 1.4, 0.7, 0.0),
             (24, 16, 32, 4, 0.3, 1.9, 1.2, 0.0), (16, 16, 16, 3, 1.0, 1.0, 1.0, 0.2)]
    prms = [pf.make_params(t_step=0.005, grad_coeff_phi=k, grad_coeff_c=k, D_L=m, D_S=m, mob_phi=l, D_gb=d)
            for (_, _, _, _, k, m, l, d) in cases]
    ics_ = [pf.ic_voronoi(nx, ny, nz, var, 5, -(i + 1)) for i, (nx, ny, nz, var, *_r) in enumerate(cases)]
    alone = []
    for (nx, ny, nz, var, *_r), prm, (phi, con) in zip(cases, prms, ics_):
        with pf.Simulation(prm, nx, ny, nz, var, "fp32_mixed") as s:
            s.put_state(phi, con)
            s.step(9)
            alone.append((s.get_state(), s.diagnose()))
    sims = []
    try:
        for (nx, ny, nz, var, *_r), prm, (phi, con) in zip(cases, prms, ics_):
            s = pf.Simulation(prm, nx, ny, nz, var, "fp32_mixed")
            s.put_state(phi, con)
            sims.append(s)
        with pf.Simulation(pf.make_params(**VOR), 16, 16, 16, 3, "fp64_faithful") as f64:
            f64.put_state(*ics_[0])
            pf.ensemble_step(sims + [f64], 4)
            pf.ensemble_step(sims + [f64], 5)
        for s, ((pa, ca), da) in zip(sims, alone):
            pb, cb = s.get_state()
            assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
            db = s.diagnose()
            assert db["step"] == da["step"] == 10 and db["e_tot"] == da["e_tot"]
    finally:
        for s in sims:
            s.close()


# Every supported axis length through the fused kernels (x lines, y panels, z
# columns), small enough in the other axes for the oracle.
AXIS_SHAPES = [(8, 12, 16), (12, 8, 24), (48, 16, 8), (96, 8, 12), (160, 8, 8), (192, 12, 8),
               (64, 48, 8), (8, 96, 12), (8, 12, 160), (16, 192, 8), (128, 8, 12), (8, 8, 256)]


@pytest.mark.parametrize("shape", AXIS_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("mode,tol_s,tol_e", [("fp64_reduced", 1e-13, 1e-13), ("fp32_mixed", 2e-6, 2e-6)])
def test_axis_lengths_vs_oracle(pf, oracle, shape, mode, tol_s, tol_e):
    _compare(pf, oracle, shape, 4, 6, 4, mode, dict(VOR, D_gb=0.25), tol_s, tol_e)


@pytest.mark.parametrize("shape", [(160, 8, 8), (12, 96, 8), (8, 12, 192)], ids=lambda s: "x".join(map(str, s)))
def test_axis_lengths_faithful_vs_oracle(pf, oracle, shape):
    _compare(pf, oracle, shape, 3, 5, 4, "fp64_faithful", dict(VOR, D_gb=0.25), 1e-13, 1e-13)


# One line per x-line block with fields = 1 mod 64 / T (T threads per line: 8 at
# 64 and 128 points, 16 at 256): the block's last wave holds one job, which its
# 64 lanes transform together with the radix-2 / radix-4 plan.  The mobility
# chain's one-field blocks (D_gb != 0) take that plan as well.
LONE_JOB = [((128, 8, 8), 8), ((256, 8, 8), 8), ((256, 8, 8), 16), ((512, 8, 8), 4), ((256, 8, 8), 4)]


@pytest.mark.parametrize("shape,var", LONE_JOB, ids=lambda s: "x".join(map(str, s)) if isinstance(s, tuple) else str(s))
@pytest.mark.parametrize("mode,tol_s,tol_e", [("fp64_reduced", 1e-13, 1e-13), ("fp32_mixed", 2e-6, 2e-6)])
def test_lone_xline_job_vs_oracle(pf, oracle, shape, var, mode, tol_s, tol_e):
    _compare(pf, oracle, shape, var, 6, 4, mode, dict(VOR, D_gb=0.25), tol_s, tol_e, tuning=dict(xline_lines=1))
