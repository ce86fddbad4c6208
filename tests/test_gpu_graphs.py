"""Graph replay of the step (pf_set_graphs): a handle that replays captured
step periods must give bit-identical states and diagnostics rows to one that
launches every kernel, for every form, diagnostics pattern (none, every step,
every 2nd, every 3rd with calls that start mid-period) and the
variable-mobility path.  Each case makes several calls in a row with state
read-back (host synchronisation and eager copies) between them, so each
captured graph is replayed many times, interleaved with eager work, and every
replay's output is checked."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = dict(t_step=0.005, grad_coeff_phi=1.0, grad_coeff_c=1.0)


def _pair(pf, mode, shape, var, params):
    nx, ny, nz = shape
    phi, con = pf.ic_voronoi(nx, ny, nz, var, 6)
    sims = []
    for graphs in (True, False):
        s = pf.Simulation(pf.make_params(**params), nx, ny, nz, var, mode)
        s.set_graphs(graphs)
        s.put_state(phi, con)
        sims.append(s)
    return sims


def _same_rows(a, b):
    assert [d["step"] for d in a] == [d["step"] for d in b]
    for da, db in zip(a, b):
        for k in da:
            assert da[k] == db[k] or (np.isnan(da[k]) and np.isnan(db[k])), (k, da, db)


@pytest.mark.parametrize("mode", ["fp32_mixed", "fp64_reduced", "fp64_faithful"])
@pytest.mark.parametrize("diag_every", [0, 1, 2, 3])
def test_graph_replay_bit_identical(pf, mode, diag_every):
    g, e = _pair(pf, mode, (16, 24, 32), 4, P)
    try:
        for n in (7, 5, 9, 2, 1, 6):
            ra, rb = g.step(n, diag_every=diag_every), e.step(n, diag_every=diag_every)
            _same_rows(ra, rb)
            pa, ca = g.get_state()
            pb, cb = e.get_state()
            assert np.array_equal(pa, pb) and np.array_equal(ca, cb), (mode, diag_every, n)
    finally:
        g.close()
        e.close()


@pytest.mark.parametrize("mode", ["fp32_mixed", "fp64_reduced"])
def test_graph_replay_variable_mobility(pf, mode):
    g, e = _pair(pf, mode, (32, 16, 16), 3, dict(P, D_gb=0.3))
    try:
        for n in (8, 3):
            _same_rows(g.step(n, diag_every=2), e.step(n, diag_every=2))
            pa, ca = g.get_state()
            pb, cb = e.get_state()
            assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
    finally:
        g.close()
        e.close()


def test_graph_replay_after_row_buffer_growth(pf):
    """A call longer than any before reallocates the diagnostics rows; the
    captured period that wrote into the old rows must not be replayed."""
    g, e = _pair(pf, "fp32_mixed", (16, 16, 16), 3, P)
    try:
        for n in (6, 200, 4):
            _same_rows(g.step(n, diag_every=2), e.step(n, diag_every=2))
        pa, ca = g.get_state()
        pb, cb = e.get_state()
        assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
    finally:
        g.close()
        e.close()


def test_ensemble_graphs_bit_identical(pf):
    shapes = [(16, 16, 16), (24, 16, 32), (32, 32, 32)]
    runs = []
    for graphs in (True, False):
        sims = []
        for (nx, ny, nz) in shapes:
            phi, con = pf.ic_voronoi(nx, ny, nz, 3, 5)
            s = pf.Simulation(pf.make_params(**P), nx, ny, nz, 3, "fp32_mixed")
            s.set_graphs(graphs)
            s.put_state(phi, con)
            sims.append(s)
        pf.ensemble_step(sims, 6)
        pf.ensemble_step(sims, 3)
        runs.append([s.get_state() for s in sims])
        for s in sims:
            s.close()
    for (pa, ca), (pb, cb) in zip(*runs):
        assert np.array_equal(pa, pb) and np.array_equal(ca, cb)
