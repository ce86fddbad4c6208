"""Checks that pin the parts of the CPU oracle the closed forms of
test_oracle.py leave open, each derived without the oracle's own formulas:

* the multi-field step (var > 1, loops 5-9 and 14 composed over fields): a
  spatially uniform state has no k != 0 modes, so one step is the explicit
  update phi_i' = phi_i - L dt df/dphi_i with c unchanged; df/dphi_i is the
  complex-step derivative of the free-energy density (src/var_diff.f90:685-759),
  not the derivative formulas of loops 7-9;
* the same composition at k != 0: the step linearised about a uniform state
  (Hessian by differences of that gradient) applied to a cosine
  mode in every field, against the oracle's step at +-eps;
* the matrix-grain symmetry: permuting eta_1..eta_{V-1} permutes the result,
  and moving the precipitate field (the last, src/var_diff.f90:738,829-835) does not;
* the extrema of src/var_diff.f90:1076-1094 against numpy, with the reference's
  initial values (phi_min 1, phi_3 0, c_max 0, c_min 1), its comparison
  operators (.le. for c_min; signed zeros tell .le. from .lt.) and phi(var);
* the summary-row timing (Q7): the row of step s carries the energies and the
  extrema of the state at the start of step s, computed here with numpy.fft;
* the initial-condition helpers of oracle/ics.py: ellipsoid masks against
  lattice-point counts and the insertion rejection rules on a hand-built state.
"""
import numpy as np
import pytest

import oracle
from oracle import ics

PAR = dict(t_step=0.01, grad_coeff_phi=0.8, grad_coeff_c=0.6, D_L=1.2, D_gb=0.6, mob_phi=0.9,
           sigma_1=1.3, sigma_2=0.7, am=0.5, ap=5.0, con_0_mat=0.1, con_0_ppt=1.0)


# ---------------------------------------------------------------------------
# the free-energy density from its definition (no derivative formulas)
# ---------------------------------------------------------------------------
def gamma_def(V, s1, s2):
    """gamma_ij: s1 between two distinct matrix grains, s2 between a grain and
    the precipitate field V-1, 0 on the diagonal (src/var_diff.f90:206-218)."""
    g = np.full((V, V), float(s1))
    g[V - 1, :] = s2
    g[:, V - 1] = s2
    np.fill_diagonal(g, 0.0)
    return g


def density(u, p):
    """(f_ch, f_eta) per point; u = (..., F) with u[..., 0] = c, u[..., 1:] = eta."""
    c, phi = u[..., 0], u[..., 1:]
    V = phi.shape[-1]
    g = gamma_def(V, p["sigma_1"], p["sigma_2"])
    a = phi * phi
    f_eta = (-0.5 * a + 0.25 * a * a).sum(-1)
    for i in range(V):
        for j in range(i + 1, V):
            f_eta = f_eta + g[i, j] * a[..., i] * a[..., j]
    d4 = (a * a).sum(-1)
    h = (a[..., : V - 1] ** 2).sum(-1) / d4
    fm = p["am"] * (c - p["con_0_mat"]) ** 2
    fp = p["ap"] * (c - p["con_0_ppt"]) ** 2
    return h * fm + (1 - h) * fp, f_eta


def grad_fd(u, p, hh=1e-30):
    """d(f_ch + f_eta)/du_f by the complex-step derivative Im f(u + i h e_f) / h:
    the density is a rational function, so this is exact to rounding (no
    cancellation, unlike a difference quotient)."""
    u = np.asarray(u, float)
    out = np.zeros_like(u)
    for k in range(u.size):
        z = u.astype(complex)
        z[k] += 1j * hh
        out[k] = sum(density(z, p)).imag / hh
    return out


def hess_fd(u, p, hh=1e-3):
    u = np.asarray(u, float)
    n = u.size
    H = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = hh
        H[:, k] = (-grad_fd(u + 2 * e, p) + 8 * grad_fd(u + e, p) - 8 * grad_fd(u - e, p)
                   + grad_fd(u - 2 * e, p)) / (12 * hh)
    return 0.5 * (H + H.T)


def to_state(u_fields):
    """(nz, ny, nx, F) with field 0 = c -> oracle arrays (phi, con)."""
    return np.ascontiguousarray(u_fields[..., 1:]), np.ascontiguousarray(u_fields[..., 0])


# ---------------------------------------------------------------------------
# uniform multi-field state
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("u0", [[0.37, 0.62, 0.35, -0.2, 0.45],
                                [0.81, 0.1, 0.7, 0.3, -0.5, 0.25, 0.6]])
def test_uniform_multifield_step_is_explicit_energy_gradient_update(u0):
    p = dict(PAR)
    shape = (4, 6, 8)  # nz, ny, nx
    P = np.prod(shape)
    u = np.array(u0, float)
    for step in range(2):
        phi, con = to_state(np.broadcast_to(u, shape + (u.size,)).copy())
        rows = oracle.run(oracle.make_params(**p), phi, con, 1)
        psi = grad_fd(u, p)
        expect = u[1:] - p["mob_phi"] * p["t_step"] * psi[1:]
        assert np.abs(phi - expect).max() < 1e-13, step
        assert np.abs(con - u[0]).max() < 1e-15  # k = 0 of c is invariant
        fch, feta = density(u, p)
        assert rows[0]["f_ch"] == pytest.approx(P * fch, rel=1e-13)
        assert rows[0]["f_eta"] == pytest.approx(P * feta, rel=1e-13)
        assert abs(rows[0]["f_int"]) < 1e-20
        # the increment is far above the tolerances: a dropped term would show
        assert np.abs(expect - u[1:]).min() > 1e-5
        u = np.concatenate([[u[0]], expect])


# ---------------------------------------------------------------------------
# linearised step at k != 0, all fields coupled, D_gb != 0
# ---------------------------------------------------------------------------
def _axis(m, n, zero_nyq):
    if zero_nyq and 2 * m == n:
        return 0.0
    return (m if m <= n // 2 else m - n) * 2 * np.pi / n


def linear_step(p, ubar, d, mode, nxyz):
    """The semi-implicit step (the update of src/var_diff.f90:967-1050 in exact
    arithmetic) applied to the cosine-mode perturbation d (an F-vector) of the
    uniform state ubar.  For a cosine mode every coefficient below is even in k,
    so the new perturbation is the same mode times the returned F-vector."""
    a = [_axis(m, n, False) for m, n in zip(mode, nxyz)]
    ap = [_axis(m, n, True) for m, n in zip(mode, nxyz)]
    s = sum(ap)                                   # S1 = i s
    q = sum(x * y * y for x, y in zip(ap, a))     # K3 = -i q
    k2 = -sum(x * x for x in a)
    k4 = sum(x ** 4 for x in a)
    H = hess_fd(ubar, p)
    lin = H @ d                                   # perturbation of (mu_c, psi_i)
    dt, L, D = p["t_step"], p["mob_phi"], p["D_L"]
    mu0 = 2 * (1 - (ubar[1:] ** 2).sum()) * p["D_gb"]   # m - D_mean of the uniform state
    out = np.zeros_like(d)
    # c: (c^ + dt S1 mu0 (S1 f^ - kc c^ K3) + D K2 dt f^) / (1 + D dt kc K4)
    out[0] = (d[0] + dt * mu0 * (-s * s * lin[0] - p["grad_coeff_c"] * s * q * d[0])
              + D * k2 * dt * lin[0]) / (1 + D * dt * p["grad_coeff_c"] * k4)
    # eta_i: (eta^ - L dt psi^) / (1 - L kphi dt K2)
    out[1:] = (d[1:] - L * dt * lin[1:]) / (1 - L * p["grad_coeff_phi"] * dt * k2)
    return out


@pytest.mark.parametrize("mode", [(1, 1, 2), (4, 0, 3), (3, 3, 5)])
def test_linearised_multifield_mode(mode):
    p = dict(PAR)
    nxyz = (8, 6, 10)
    ubar = np.array([0.42, 0.55, 0.3, 0.2, 0.4])
    d = np.array([0.3, -0.5, 0.8, 0.1, -0.4])
    nx, ny, nz = nxyz
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    th = 2 * np.pi * (mode[0] * x / nx + mode[1] * y / ny + mode[2] * z / nz)
    eps = 1e-5  # antisymmetric difference: the residual is O(eps^2), measured <= 4e-8 relative
    res = []
    for sgn in (1.0, -1.0):
        u = ubar + sgn * eps * d * np.cos(th)[..., None]
        phi, con = to_state(u)
        oracle.run(oracle.make_params(**p), phi, con, 1)
        res.append(np.concatenate([con[..., None], phi], axis=-1))
    got = (res[0] - res[1]) / (2 * eps)
    g = linear_step(p, ubar, d, mode, nxyz)
    expect = g * np.cos(th)[..., None]
    inc = np.abs(g - d)  # each field's increment over the step
    assert inc.min() > 1e-4  # every field moves
    err = np.abs(got - expect).reshape(-1, d.size).max(axis=0)
    assert np.all(err < 1e-6 * inc), err / inc


# ---------------------------------------------------------------------------
# matrix-grain symmetry
# ---------------------------------------------------------------------------
def _noisy_polycrystal(n, var, seeds, seed):
    phi, con = oracle.ic_voronoi(n, n, n, var, seeds)
    rng = np.random.default_rng(seed)
    phi += 0.05 * rng.standard_normal(phi.shape)
    return phi, con


def test_matrix_grain_permutation_permutes_the_step():
    p = oracle.make_params(**dict(PAR, t_step=0.005))
    var = 5
    phi, con = _noisy_polycrystal(8, var, 6, 3)
    perm = [2, 0, 3, 1, 4]  # matrix grains permuted, precipitate (last) fixed
    a_phi, a_con = phi.copy(), con.copy()
    b_phi, b_con = np.ascontiguousarray(phi[..., perm]), con.copy()
    ra = oracle.run(p, a_phi, a_con, 3)
    rb = oracle.run(p, b_phi, b_con, 3)
    assert np.abs(b_phi - a_phi[..., perm]).max() < 1e-13
    assert np.abs(b_con - a_con).max() < 1e-13
    for x, y in zip(ra, rb):
        for k in oracle.DIAG_FIELDS:
            assert y[k] == pytest.approx(x[k], rel=1e-12, abs=1e-14)
    # the precipitate field is not interchangeable with a grain
    swap = [4, 1, 2, 3, 0]
    c_phi, c_con = np.ascontiguousarray(phi[..., swap]), con.copy()
    oracle.run(p, c_phi, c_con, 3)
    assert np.abs(c_phi - a_phi[..., swap]).max() > 1e-4


# ---------------------------------------------------------------------------
# extrema
# ---------------------------------------------------------------------------
def np_extrema(phi, con):
    """Numpy restatement for data without signed zeros: the reference's initial
    values and phi(var) for phi_3."""
    sq = np.zeros(con.shape)
    for i in range(phi.shape[-1]):
        sq = sq + phi[..., i] * phi[..., i]
    return dict(c_max=max(0.0, con.max()), c_min=min(1.0, con.min()),
                phi_min=min(1.0, sq.min()), phi_3_max=max(0.0, phi[..., -1].max()))


@pytest.mark.parametrize("case", ["random", "c_above_one", "c_negative", "phi_sq_above_one",
                                  "precipitate_negative", "ties"])
def test_extrema_vs_numpy(case):
    rng = np.random.default_rng(11)
    phi = rng.uniform(-0.5, 0.9, (6, 5, 4, 5))
    con = rng.uniform(0.1, 0.9, (6, 5, 4))
    phi[..., 2] += 0.6  # a large third field: phi_3 must still be the last field's
    if case == "c_above_one":
        con += 1.5
    elif case == "c_negative":
        con -= 2.0
    elif case == "phi_sq_above_one":
        phi[..., 0] = 1.2
    elif case == "precipitate_negative":
        phi[..., -1] = -np.abs(phi[..., -1]) - 0.01
    elif case == "ties":
        con[::2] = con.max()
        con[1::2] = con.min()
        phi[..., -1] = 0.3
    got = oracle.extrema(phi, con)
    want = np_extrema(phi, con)
    for k in want:
        assert got[k] == want[k], k
    if case == "c_above_one":
        assert got["c_min"] == 1.0
    if case == "c_negative":
        assert got["c_max"] == 0.0
    if case == "phi_sq_above_one":
        assert got["phi_min"] == 1.0
    if case == "precipitate_negative":
        assert got["phi_3_max"] == 0.0


def test_extrema_comparison_operators_signed_zero():
    """c_min uses .le. (a later equal value replaces it), c_max and phi_3 use
    .gt. from +0 (src/var_diff.f90:1081-1083): signed zeros show which."""
    phi = np.zeros((1, 1, 2, 2))
    phi[..., 0] = 0.5
    phi[0, 0, :, 1] = [-0.0, -0.0]
    got = oracle.extrema(phi, np.array([[[0.0, -0.0]]]))
    assert got["c_min"] == 0.0 and np.signbit(got["c_min"])        # .le.: -0 replaced +0
    assert got["c_max"] == 0.0 and not np.signbit(got["c_max"])    # .gt.: -0 never replaces +0
    assert not np.signbit(got["phi_3_max"])
    got = oracle.extrema(phi, np.array([[[-0.0, 0.0]]]))
    assert got["c_min"] == 0.0 and not np.signbit(got["c_min"])


# ---------------------------------------------------------------------------
# summary-row timing and contents (Q7)
# ---------------------------------------------------------------------------
def np_fint(phi, con, p):
    """f_int = sum_r 1/2 kappa |grad' u|^2 by numpy.fft with the Nyquist-zeroed
    wavenumbers (src/k_space.f90:51-59)."""
    nz, ny, nx = con.shape
    kx = np.array([_axis(m, nx, True) for m in range(nx // 2 + 1)])
    ky = np.array([_axis(m, ny, True) for m in range(ny)])
    kz = np.array([_axis(m, nz, True) for m in range(nz)])
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    tot = 0.0
    fields = [(con, p["grad_coeff_c"])] + [(phi[..., i], p["grad_coeff_phi"]) for i in range(phi.shape[-1])]
    for u, kap in fields:
        U = np.fft.rfftn(u)
        g2 = sum(np.fft.irfftn(1j * K * U, s=u.shape, axes=(0, 1, 2)) ** 2 for K in (KX, KY, KZ))
        tot += 0.5 * kap * g2.sum()
    return tot


def test_row_of_step_s_describes_the_state_at_its_start():
    p = dict(PAR, t_step=0.005, D_gb=0.3)
    P_ = oracle.make_params(**p)
    phi, con = _noisy_polycrystal(8, 4, 5, 7)
    rows = oracle.run(P_, phi.copy(), con.copy(), 4)
    states = []
    a_phi, a_con = phi.copy(), con.copy()
    for s in range(4):
        states.append((a_phi.copy(), a_con.copy()))
        oracle.run(P_, a_phi, a_con, 1)
    for s, (sp, sc) in enumerate(states):
        want = np_extrema(sp, sc)
        for k in want:
            assert rows[s][k] == want[k], (s, k)
        fch, feta = density(np.concatenate([sc[..., None], sp], axis=-1), p)
        assert rows[s]["f_ch"] == pytest.approx(fch.sum(), rel=1e-12)
        assert rows[s]["f_eta"] == pytest.approx(feta.sum(), rel=1e-12)
        assert rows[s]["f_int"] == pytest.approx(np_fint(sp, sc, p), rel=1e-11)
        assert rows[s]["e_tot"] == (rows[s]["f_int"] + rows[s]["f_ch"]) + rows[s]["f_eta"]
    # the states differ from step to step, so a row shifted by one would fail
    assert np_extrema(*states[0])["phi_min"] != np_extrema(*states[1])["phi_min"]


# ---------------------------------------------------------------------------
# initial conditions and insertion (oracle/ics.py)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("rad", [(3, 3, 3), (2, 4, 3), (5, 1, 2), (4, 4, 4)])
def test_ellipsoid_mask_against_lattice_points(rad):
    offs = set(ics._ellipsoid_offsets(rad))
    rx, ry, rz = rad
    inside = boundary = 0
    for k in range(-rz, rz + 1):
        for j in range(-ry, ry + 1):
            for i in range(-rx, rx + 1):
                # (i/rx)^2 + (j/ry)^2 + (k/rz)^2 against 1 in exact integers
                lhs = i * i * ry * ry * rz * rz + j * j * rx * rx * rz * rz + k * k * rx * rx * ry * ry
                rhs = rx * rx * ry * ry * rz * rz
                if lhs < rhs:
                    inside += 1
                    assert (i, j, k) in offs
                elif lhs > rhs:
                    assert (i, j, k) not in offs
                else:
                    boundary += 1
    # points on the surface: the axis tips (i/r = 1 exactly in single precision) are in
    for t in ((rx, 0, 0), (0, ry, 0), (0, 0, rz), (-rx, 0, 0)):
        assert t in offs
    assert inside <= len(offs) <= inside + boundary


def test_circle_grain_is_the_lattice_ball():
    n, g = 16, 4
    cfg = dict(nx=n, ny=n, nz=n, var=3, grn_rad=g, ppt_rad=(2, 2, 2), num_ppt=0, con_0_ppt=1.0)
    phi = np.zeros((n, n, n, 3))
    con = np.full((n, n, n), 0.1)
    ics.circle(cfg, phi, con, oracle.Ran2(-1))
    k, j, i = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    r2 = (i - n // 2) ** 2 + (j - n // 2) ** 2 + (k - n // 2) ** 2
    grain = phi[..., 1] == 1.0
    assert np.all(grain[r2 < g * g]) and not np.any(grain[r2 > g * g])
    assert np.all(phi[..., 0][~grain] == 1.0) and np.all(phi[..., 0][grain] == 0.0)
    assert np.all(phi[..., 2] == 0.0)  # no precipitates requested


class _Seq:
    """A ran_2 stand-in returning fixed values in order."""

    def __init__(self, vals):
        self.vals = list(vals)
        self.n = 0

    def __call__(self):
        self.n += 1
        return self.vals.pop(0)


def _u(idx, n):
    # ix = int(u*N + 1.0) = idx for u in the middle of its bin
    return (idx - 0.5) / n


def test_insertion_rejection_rules():
    n, var = 12, 3
    cfg = dict(nx=n, ny=n, nz=n, var=var, ppt_rad=(1, 1, 1), num_ppt=1)
    phi = np.zeros((n, n, n, var))
    phi[..., 0] = 0.5               # sum of matrix phi^2 = 0.25: admissible
    phi[2, 2, 2, 0] = 0.95          # centre (3,3,3): 0.9025 > 0.8, rejected
    phi[7, 7, 7, var - 1] = 0.01    # precipitate at (8,8,8): a centre within rad of it is rejected
    phi[4, 9, 9, 0] = np.sqrt(0.8)  # centre (10,10,5): exactly 0.8 in double, not > 0.8 in single: admissible
    draws = [_u(3, n), _u(3, n), _u(3, n),     # rejected: matrix sum
             _u(8, n), _u(9, n), _u(8, n),     # rejected: overlaps (8,8,8) (distance^2 = 1 <= 3)
             _u(10, n), _u(10, n), _u(5, n)]   # accepted
    rng = _Seq(draws)
    before = phi.copy()
    ics.insert(cfg, phi, rng)
    assert rng.n == 9
    k, j, i = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    ball = (i - 10) ** 2 + (j - 10) ** 2 + (k - 5) ** 2 <= 3
    assert ball.sum() == 1 + 6 + 12 + 8  # lattice points within sqrt(3): centre, faces, edges, corners
    assert np.all(phi[ball, var - 1] == 1.0) and np.all(phi[ball, : var - 1] == 0.0)
    assert np.array_equal(phi[~ball], before[~ball])
