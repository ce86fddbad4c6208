"""Python host binding of libpf.so, the MI355X phase-field timestep library.

The names mirror the reference's Fortran interface for this path
(src/var_diff.f90 time loop, src/f_trans.f90, src/inv_trans.f90,
src/k_space.f90, src/ran_2.f90); arrays use the reference's layouts expressed as
C-ordered numpy arrays:

    phi   (nz, ny, nx, var)      == Fortran phi(var, nx, ny, nz)
    con   (nz, ny, nx)           == Fortran con(nx, ny, nz)
    spectrum (nx//2+1, ny, nz)   == Fortran (nz, ny, nx/2+1), kz fastest (P3DFFT Z-pencil)

The library is required: importing this module raises if libpf.so has not been
built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpf.so")

PF_OK, PF_EINVAL, PF_ENOMEM, PF_EHIP, PF_ENOTIMPL, PF_ENONFINITE = 0, -1, -2, -3, -4, -5
PREC_FP64, PREC_FP32 = 0, 1
LAYOUT_XFAST, LAYOUT_ZFAST = 0, 1
MODE_FP64_FAITHFUL, MODE_FP64_REDUCED, MODE_FP32_MIXED = 0, 1, 2
MODES = {"fp64_faithful": 0, "fp64_reduced": 1, "fp32_mixed": 2}
SUPPORTED_SIZES = (8, 12, 16, 24, 32, 48, 64, 96, 128, 160, 192, 256, 512)

PARAM_NAMES = ("t_step", "grad_coeff_phi", "grad_coeff_c", "D_S", "D_L", "D_gb", "mob_phi",
               "sigma_1", "sigma_2", "am", "ap", "con_0_mat", "con_0_ppt")
DIAG_NAMES = ("f_int", "f_ch", "f_eta", "e_tot", "c_max", "c_min", "phi_min", "phi_3_max", "c_mean")


class PfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libpf error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in PARAM_NAMES]


class Diag(C.Structure):
    _fields_ = [("step", C.c_int), ("valid", C.c_int)] + [(n, C.c_double) for n in DIAG_NAMES]

    def as_dict(self):
        d = {n: getattr(self, n) for n in DIAG_NAMES}
        d["step"] = self.step
        d["valid"] = bool(self.valid)
        return d


class Ran2State(C.Structure):
    _fields_ = [("idum2", C.c_int), ("iy", C.c_int), ("iv", C.c_int * 32)]


TUNING_NAMES = ("xline_lines", "xline_grid", "zline_grid", "zline_run_rows", "zline_plan")
TUNING_DEFAULTS = dict(xline_lines=0, xline_grid=0, zline_grid=0, zline_run_rows=-1, zline_plan=0)


class Tuning(C.Structure):
    _fields_ = [(n, C.c_int) for n in TUNING_NAMES]


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

EXPORTED = (
    "pf_create", "pf_destroy", "pf_put_state", "pf_get_state", "pf_step", "pf_diagnose",
    "pf_set_timing", "pf_stage_times", "pf_stage_name", "pf_stage_bytes", "pf_stream",
    "pf_fft_r2c", "pf_fft_c2r", "pf_f_trans", "pf_inv_trans", "pf_ran2_init", "pf_ran2",
    "pf_k_space", "pf_gamma", "pf_ic_random", "pf_ic_voronoi", "pf_last_error", "pf_version",
    "pf_ensemble_step", "pf_analyze_state", "pf_set_graphs", "pf_put_voronoi", "pf_fft_r2c_noalloc",
    "pf_set_tuning", "pf_get_tuning", "pf_set_timing_stages",
)


def build() -> None:
    """Compile libpf.so in-tree (hipcc, gfx950)."""
    subprocess.run(["make", "-C", HERE, "-j8", "libpf.so"], check=True)


def lib():
    """Load libpf.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                          f"or `make -C {HERE}`")
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64 (same
    # soname as /opt/rocm's).  Loaded first, it also satisfies libpf.so's
    # dependency; loaded after libpf.so, a second runtime would be mapped and
    # torch.cuda would find no device.  So torch, when installed, goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    L.pf_create.argtypes = [C.POINTER(Params), ip, ip, ip, ip, ip, C.POINTER(vp)]
    L.pf_destroy.argtypes = [vp]
    L.pf_put_state.argtypes = [vp, _dp, _dp]
    L.pf_get_state.argtypes = [vp, _dp, _dp]
    L.pf_put_voronoi.argtypes = [vp, ip, ip, dp, dp]
    L.pf_step.argtypes = [vp, ip, ip, vp]
    L.pf_diagnose.argtypes = [vp, C.POINTER(Diag)]
    L.pf_set_timing.argtypes = [vp, ip]
    L.pf_set_graphs.argtypes = [vp, ip]
    L.pf_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.pf_set_timing_stages.argtypes = [vp, C.c_uint]
    L.pf_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.pf_stage_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_long), ip]
    L.pf_stage_name.argtypes = [ip, C.c_char_p, ip]
    L.pf_stage_bytes.argtypes = [vp, ip]
    L.pf_stage_bytes.restype = C.c_double
    L.pf_stream.argtypes = [vp]
    L.pf_stream.restype = vp
    L.pf_fft_r2c.argtypes = [ip, ip, ip, ip, ip, vp, vp, ip, dp, vp]
    L.pf_fft_c2r.argtypes = [ip, ip, ip, ip, ip, vp, vp, ip, vp]
    L.pf_f_trans.argtypes = [_dp, _dp, ip, ip, ip]
    L.pf_inv_trans.argtypes = [_dp, _dp, ip, ip, ip]
    L.pf_ran2_init.argtypes = [C.POINTER(Ran2State)]
    L.pf_ran2_init.restype = None
    L.pf_ran2.argtypes = [C.POINTER(Ran2State), C.POINTER(C.c_int)]
    L.pf_ran2.restype = C.c_double
    L.pf_k_space.argtypes = [ip, ip, ip, _dp, _dp, _dp]
    L.pf_gamma.argtypes = [ip, dp, dp, _dp]
    L.pf_ic_random.argtypes = [ip, ip, ip, ip, ip, dp, dp, dp, _dp, _dp]
    L.pf_ic_voronoi.argtypes = [ip, ip, ip, ip, ip, ip, dp, dp, _dp, _dp]
    L.pf_last_error.restype = C.c_char_p
    L.pf_ensemble_step.argtypes = [C.POINTER(vp), ip, ip]
    L.pf_analyze_state.argtypes = [_dp, ip, ip, ip, ip, dp, _dp, _dp, _dp, _dp,
                                   np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS"),
                                   np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
    _lib = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        raise PfError(rc, lib().pf_last_error().decode())
    return rc


def make_params(**kw) -> Params:
    """Parameters named as in input.txt; defaults = inputs/config.yaml's free energy."""
    d = dict(t_step=0.01, grad_coeff_phi=0.5, grad_coeff_c=0.5, D_S=1.0, D_L=1.0, D_gb=0.0,
             mob_phi=1.0, sigma_1=1.5, sigma_2=1.5, am=0.5, ap=5.0, con_0_mat=0.1,
             con_0_ppt=1.0)
    unknown = set(kw) - set(d)
    if unknown:
        raise TypeError(f"unknown parameters {sorted(unknown)}")
    d.update(kw)
    return Params(**d)


# ---------------------------------------------------------------------------
# Timestep
# ---------------------------------------------------------------------------
class Simulation:
    """One grid on the GPU: replaces the time-loop body of src/var_diff.f90:474-1156."""

    def __init__(self, params: Params, nx: int, ny: int, nz: int, var: int,
                 mode: str | int = "fp64_faithful"):
        m = MODES[mode] if isinstance(mode, str) else int(mode)
        self.nx, self.ny, self.nz, self.var, self.mode = nx, ny, nz, var, m
        self.params = params
        h = C.c_void_p()
        _check(lib().pf_create(C.byref(params), nx, ny, nz, var, m, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().pf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def put_state(self, phi: np.ndarray, con: np.ndarray) -> None:
        assert phi.shape == (self.nz, self.ny, self.nx, self.var), phi.shape
        assert con.shape == (self.nz, self.ny, self.nx), con.shape
        _check(lib().pf_put_state(self._h, np.ascontiguousarray(phi, dtype=np.float64),
                                  np.ascontiguousarray(con, dtype=np.float64)))

    def put_voronoi(self, nseeds: int, iseed: int = -1, c0: float = 0.4, c_amp: float = 0.02) -> None:
        """The state ic_voronoi(...) would give, built on the device (pf_put_voronoi)."""
        _check(lib().pf_put_voronoi(self._h, nseeds, iseed, c0, c_amp))

    def get_state(self):
        phi = np.empty((self.nz, self.ny, self.nx, self.var))
        con = np.empty((self.nz, self.ny, self.nx))
        _check(lib().pf_get_state(self._h, phi, con))
        return phi, con

    def step(self, nsteps: int, diag_every: int = 0, check: bool = True):
        """Advance nsteps; returns a list of per-step diagnostics dicts (start-of-step
        state) for the steps that are multiples of diag_every."""
        out = (Diag * max(nsteps, 1))() if diag_every > 0 else None
        rc = lib().pf_step(self._h, nsteps, diag_every, C.cast(out, C.c_void_p) if out else None)
        if check:
            _check(rc)
        if out is None:
            return []
        return [out[i].as_dict() for i in range(nsteps) if out[i].valid]

    @property
    def handle(self):
        return self._h

    def diagnose(self) -> dict:
        d = Diag()
        _check(lib().pf_diagnose(self._h, C.byref(d)))
        return d.as_dict()

    def set_timing(self, on: bool) -> None:
        _check(lib().pf_set_timing(self._h, int(on)))

    def set_timing_stages(self, names) -> None:
        """Timing events for the named stages only (names as stage_times reports them)."""
        mask = 0
        for i in range(32):
            buf = C.create_string_buffer(64)
            if lib().pf_stage_name(i, buf, 64) < 0:
                break
            if buf.value.decode() in names:
                mask |= 1 << i
        _check(lib().pf_set_timing_stages(self._h, mask))

    def set_graphs(self, on: bool) -> None:
        """Replay captured step periods (hipGraph) or launch every kernel."""
        _check(lib().pf_set_graphs(self._h, int(on)))

    def set_tuning(self, **kw) -> None:
        """Launch shapes of the fused kernels (pf_set_tuning): xline_lines,
        xline_grid, zline_grid, zline_run_rows, zline_plan; unnamed ones take
        their defaults."""
        unknown = set(kw) - set(TUNING_NAMES)
        if unknown:
            raise TypeError(f"unknown tuning fields {sorted(unknown)}")
        t = Tuning(**dict(TUNING_DEFAULTS, **kw))
        _check(lib().pf_set_tuning(self._h, C.byref(t)))

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(lib().pf_get_tuning(self._h, C.byref(t)))
        return {n: getattr(t, n) for n in TUNING_NAMES}

    def stage_times(self):
        n = 32
        ms = (C.c_double * n)()
        cnt = (C.c_long * n)()
        ns = _check(lib().pf_stage_times(self._h, ms, cnt, n))
        res = {}
        for i in range(ns):
            buf = C.create_string_buffer(64)
            lib().pf_stage_name(i, buf, 64)
            res[buf.value.decode()] = dict(ms=ms[i], count=cnt[i],
                                           bytes=lib().pf_stage_bytes(self._h, i))
        return res

    @property
    def stream(self) -> int:
        return lib().pf_stream(self._h)


# ---------------------------------------------------------------------------
# Transforms and helpers with the reference's names
# ---------------------------------------------------------------------------
def f_trans(x: np.ndarray) -> np.ndarray:
    """src/f_trans.f90: forward R2C / (Nx*Ny*Nz).  x: (nz, ny, nx) real.
    Returns the spectrum as (nx//2+1, ny, nz) complex (Fortran (nz, ny, nx/2+1))."""
    nz, ny, nx = x.shape
    out = np.empty((nx // 2 + 1, ny, nz, 2))
    _check(lib().pf_f_trans(np.ascontiguousarray(x, dtype=np.float64), out, nx, ny, nz))
    return out[..., 0] + 1j * out[..., 1]


def inv_trans(h: np.ndarray, nx: int) -> np.ndarray:
    """src/inv_trans.f90: inverse C2R, unnormalised.  h: (nx//2+1, ny, nz) complex."""
    mx, ny, nz = h.shape
    assert mx == nx // 2 + 1
    hin = np.ascontiguousarray(np.stack([h.real, h.imag], axis=-1))
    out = np.empty((nz, ny, nx))
    _check(lib().pf_inv_trans(hin, out, nx, ny, nz))
    return out


def k_space(nx: int, ny: int, nz: int):
    """src/k_space.f90: kf, kf_sq, kf_4 shaped (nx//2+1, ny, nz, 3) complex."""
    mx = nx // 2 + 1
    arrs = [np.empty((mx, ny, nz, 3, 2)) for _ in range(3)]
    _check(lib().pf_k_space(nx, ny, nz, *arrs))
    return tuple(a[..., 0] + 1j * a[..., 1] for a in arrs)


def gamma(var: int, sigma_1: float, sigma_2: float) -> np.ndarray:
    g = np.empty(var * var)
    _check(lib().pf_gamma(var, sigma_1, sigma_2, g))
    return g.reshape(var, var, order="F")


class Ran2:
    """ran_2 (src/ran_2.f90) with the generator state held by this object."""

    def __init__(self, idum: int):
        self.st = Ran2State()
        lib().pf_ran2_init(C.byref(self.st))
        self.idum = C.c_int(idum)

    def __call__(self) -> float:
        return lib().pf_ran2(C.byref(self.st), C.byref(self.idum))


def ic_random(nx, ny, nz, var, iseed=-1, c0=0.5, c_amp=0.01, eta_amp=0.001):
    phi = np.empty((nz, ny, nx, var))
    con = np.empty((nz, ny, nx))
    _check(lib().pf_ic_random(nx, ny, nz, var, iseed, c0, c_amp, eta_amp, phi, con))
    return phi, con


def ic_voronoi(nx, ny, nz, var, nseeds, iseed=-1, c0=0.4, c_amp=0.02):
    phi = np.empty((nz, ny, nx, var))
    con = np.empty((nz, ny, nx))
    _check(lib().pf_ic_voronoi(nx, ny, nz, var, nseeds, iseed, c0, c_amp, phi, con))
    return phi, con


def fft_r2c_device(nx, ny, nz, batch, prec, d_in: int, d_out: int, layout=LAYOUT_XFAST,
                   scale=1.0, stream=None) -> None:
    _check(lib().pf_fft_r2c(nx, ny, nz, batch, prec, d_in, d_out, layout, scale, stream))


def fft_c2r_device(nx, ny, nz, batch, prec, d_in: int, d_out: int, layout=LAYOUT_XFAST,
                   stream=None) -> None:
    _check(lib().pf_fft_c2r(nx, ny, nz, batch, prec, d_in, d_out, layout, stream))


def ensemble_step(sims, nsteps: int) -> None:
    """Advance independent simulations together (pf_ensemble_step): each step is
    enqueued on every simulation's stream in turn, so the GPU runs them
    concurrently; returns when all have finished."""
    hs = (C.c_void_p * len(sims))(*[s.handle for s in sims])
    _check(lib().pf_ensemble_step(hs, len(sims), nsteps))


def analyze_state(phi: np.ndarray, threshold: float) -> dict:
    """src/paraview.f90:188-313 for one state phi (nz, ny, nx, var) on the GPU:
    flat per-voxel arrays (x fastest) and (precipitate boundary voxels, boundary voxels)."""
    nz, ny, nx, var = phi.shape
    P = nx * ny * nz
    out = {k: np.empty(P) for k in ("phi_sq", "phi_3", "grad_phi_3", "boundary")}
    vn = np.empty(P, dtype=np.int32)
    counts = np.zeros(2, dtype=np.int64)
    _check(lib().pf_analyze_state(np.ascontiguousarray(phi, dtype=np.float64), nx, ny, nz, var, threshold,
                                  out["phi_sq"], out["phi_3"], out["grad_phi_3"], out["boundary"], vn, counts))
    out["var_num"] = vn
    out["counts"] = (int(counts[0]), int(counts[1]))
    return out
