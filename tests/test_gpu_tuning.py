"""Launch shapes of the fused kernels (pf_set_tuning): they decide which block
does which lines, columns and tiles, not what is computed for each, so a run
with any accepted shape ends in the same state bit for bit as the defaults,
and its summary energies (sums of per-block partials) agree to 1e-12.  The z
transforms' radix sequence (zline_plan) changes the rounding: checked against
the oracle instead."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VOR = dict(t_step=0.005, grad_coeff_phi=1.0, grad_coeff_c=1.0)

SHAPES = [dict(xline_lines=2), dict(xline_lines=4, xline_grid=256), dict(xline_lines=8, xline_grid=1024),
          dict(zline_grid=1024, zline_run_rows=0), dict(zline_grid=65536, zline_run_rows=4),
          dict(xline_lines=2, xline_grid=16384, zline_grid=2048, zline_run_rows=2)]


@pytest.mark.parametrize("mode", ["fp32_mixed", "fp64_reduced"])
@pytest.mark.parametrize("dgb", [0.0, 0.3])
def test_tuned_shapes_leave_the_state_unchanged(pf, mode, dgb):
    n, var = 32, 4
    prm = pf.make_params(**dict(VOR, D_gb=dgb))
    phi, con = pf.ic_voronoi(n, n, n, var, 8)

    def run(**kw):
        with pf.Simulation(prm, n, n, n, var, mode) as sim:
            sim.set_tuning(**kw)
            sim.put_state(phi, con)
            rows = sim.step(8, diag_every=2)
            return sim.get_state(), rows, sim.get_tuning()

    (p0, c0), r0, t0 = run()
    assert t0 == pf.TUNING_DEFAULTS
    for kw in SHAPES:
        print(mode, dgb, kw, flush=True)
        (p1, c1), r1, t1 = run(**kw)
        assert t1 == dict(pf.TUNING_DEFAULTS, **kw)
        assert np.array_equal(p0, p1) and np.array_equal(c0, c1), kw
        for a, b in zip(r0, r1):
            for k in ("f_int", "f_ch", "f_eta", "e_tot"):
                assert abs(a[k] - b[k]) <= 1e-12 * max(abs(a[k]), 1.0), (kw, k, a[k], b[k])
            for k in ("c_max", "c_min", "phi_min", "phi_3_max", "c_mean"):
                assert a[k] == b[k], (kw, k)


def test_bad_tuning_is_refused(pf):
    with pf.Simulation(pf.make_params(**VOR), 16, 16, 16, 3, "fp32_mixed") as sim:
        for kw in (dict(xline_lines=3), dict(xline_grid=100), dict(zline_grid=1 << 20),
                   dict(zline_run_rows=9)):
            with pytest.raises(pf.PfError) as e:
                sim.set_tuning(**kw)
            assert e.value.code == pf.PF_EINVAL
        # 512-point x lines with 37 fields cannot go two lines to a block (1184 threads)
    with pf.Simulation(pf.make_params(**VOR), 512, 8, 8, 36, "fp32_mixed") as sim:
        with pytest.raises(pf.PfError):
            sim.set_tuning(xline_lines=2)
        sim.set_tuning(xline_lines=1)


# The z transforms' alternative radix sequences (zline_plan 1, 2; FP32-mixed):
# different rounding, so compared with the oracle at the FP32 parity bound.
@pytest.mark.parametrize("nz", [64, 128, 160, 192, 256, 512])
@pytest.mark.parametrize("plan", [1, 2])
def test_zline_plan_variants_vs_oracle(pf, oracle, nz, plan):
    nx, ny, var = 8, 12, 4
    prm = dict(VOR, D_gb=0.25)
    phi, con = pf.ic_voronoi(nx, ny, nz, var, 6)
    with pf.Simulation(pf.make_params(**prm), nx, ny, nz, var, "fp32_mixed") as sim:
        sim.set_tuning(zline_plan=plan)
        assert sim.get_tuning()["zline_plan"] == plan
        sim.put_state(phi, con)
        rows = sim.step(6, diag_every=2)
        gphi, gcon = sim.get_state()
    ophi, ocon = phi.copy(), con.copy()
    od = oracle.run(oracle.make_params(**prm), ophi, ocon, 6)
    err = max(np.abs(gphi - ophi).max(), np.abs(gcon - ocon).max())
    e = max(abs(r["e_tot"] - od[r["step"] - 1]["e_tot"]) /
            (abs(od[r["step"] - 1]["f_int"]) + abs(od[r["step"] - 1]["f_ch"]) + abs(od[r["step"] - 1]["f_eta"]))
            for r in rows)
    print(f"z plan {plan} at nz = {nz}: |state - oracle| {err:.2e}, e_tot {e:.2e}")
    assert err < 2e-6 and e < 2e-6


def test_zline_plan_refused_where_absent(pf):
    with pf.Simulation(pf.make_params(**VOR), 16, 16, 32, 3, "fp32_mixed") as sim:
        with pytest.raises(pf.PfError):
            sim.set_tuning(zline_plan=1)  # no alternative plan for 32
    with pf.Simulation(pf.make_params(**VOR), 16, 16, 64, 3, "fp64_reduced") as sim:
        with pytest.raises(pf.PfError):
            sim.set_tuning(zline_plan=1)  # FP32 only


def test_batched_ensemble_honours_tuning(pf):
    """Members of one shape with the same tuned shapes (x-line lines, z-line
    plan) step batched; each equals the member stepped alone with that tuning."""
    n, nz, var = 16, 64, 3
    kw = dict(xline_lines=2, zline_plan=2, zline_grid=2048, zline_run_rows=0)
    prms = [pf.make_params(t_step=0.005, grad_coeff_phi=k, grad_coeff_c=k, D_L=m, D_S=m, mob_phi=l)
            for k, m, l in ((0.6, 1.1, 0.9), (1.4, 0.7, 1.6), (1.0, 1.9, 0.5))]
    ics_ = [pf.ic_voronoi(n, n, nz, var, 5, -(i + 1)) for i in range(len(prms))]
    alone = []
    for prm, (phi, con) in zip(prms, ics_):
        with pf.Simulation(prm, n, n, nz, var, "fp32_mixed") as s:
            s.set_tuning(**kw)
            s.put_state(phi, con)
            s.step(7)
            alone.append(s.get_state())
    sims = []
    try:
        for prm, (phi, con) in zip(prms, ics_):
            s = pf.Simulation(prm, n, n, nz, var, "fp32_mixed")
            s.set_tuning(**kw)
            s.put_state(phi, con)
            sims.append(s)
        pf.ensemble_step(sims, 7)
        for s, (pa, ca) in zip(sims, alone):
            pb, cb = s.get_state()
            assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
    finally:
        for s in sims:
            s.close()
