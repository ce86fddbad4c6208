"""ctypes bindings for the CPU FP64 oracle (``oracle/liboracle.so``).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, and only as the checker (or the timed
CPU baseline).  The product path never imports this module.

``ref()`` returns the reference's own ``ran_2.f90`` / ``k_space.f90`` compiled
into ``oracle/_ref/libref.so`` (see ``oracle/Makefile``), or ``None`` when that
library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_lib = None
_ref = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Ran2State(C.Structure):
    _fields_ = [("idum2", C.c_int), ("iy", C.c_int), ("iv", C.c_int * 32)]


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "t_step", "grad_coeff_phi", "grad_coeff_c", "D_S", "D_L", "D_gb", "mob_phi",
        "sigma_1", "sigma_2", "am", "ap", "con_0_mat", "con_0_ppt")]


DIAG_FIELDS = ("f_int", "f_ch", "f_eta", "e_tot", "c_max", "c_min", "phi_min", "phi_3_max")


class Diag(C.Structure):
    _fields_ = [(n, C.c_double) for n in DIAG_FIELDS]


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref/libref.so when the reference tree exists)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-C", HERE, "ref"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.ora_ran2.restype = C.c_double
        L.ora_ran2.argtypes = [C.POINTER(Ran2State), C.POINTER(C.c_int)]
        L.ora_ran2_init.argtypes = [C.POINTER(Ran2State)]
        L.ora_fft_r2c.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int]
        L.ora_fft_c2r.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.ora_fft_c2c_line.argtypes = [C.c_int, _dp, _dp, C.c_int]
        L.ora_kspace_table.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.ora_twiddles.argtypes = [C.c_int, _dp]
        L.ora_gamma.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.ora_run.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.c_int,
                              _dp, _dp, C.c_int, C.c_void_p]
        L.ora_pointwise.argtypes = [C.POINTER(Params), C.c_int, C.c_int, _dp, _dp, _dp,
                                    _dp, _dp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ora_extrema.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        L.ora_ic_random.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, _dp, _dp]
        L.ora_ic_voronoi.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_double, _dp, _dp]
        L.ora_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def ref():
    """The reference's ran_2 / k_space compiled with amdflang, or None."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        R.ran_2_.restype = C.c_double
        R.ran_2_.argtypes = [C.POINTER(C.c_int)]
        _ref = R
    return _ref


# ---------------------------------------------------------------------------
class Ran2:
    """Stateful ran_2 generator (the reference keeps this state in SAVE variables)."""

    def __init__(self, idum: int):
        self.st = Ran2State()
        lib().ora_ran2_init(C.byref(self.st))
        self.idum = C.c_int(idum)

    def __call__(self) -> float:
        return lib().ora_ran2(C.byref(self.st), C.byref(self.idum))


def fft_r2c(x: np.ndarray, scale: bool = True) -> np.ndarray:
    """x: real array shaped (nz, ny, nx) (x fastest).  Returns (nz, ny, nx//2+1) complex."""
    nz, ny, nx = x.shape
    out = np.zeros((nz, ny, nx // 2 + 1, 2))
    rc = lib().ora_fft_r2c(nx, ny, nz, np.ascontiguousarray(x, dtype=np.float64), out, int(scale))
    assert rc == 0
    return out[..., 0] + 1j * out[..., 1]


def fft_c2r(h: np.ndarray, nx: int) -> np.ndarray:
    nz, ny, mx = h.shape
    assert mx == nx // 2 + 1
    hin = np.ascontiguousarray(np.stack([h.real, h.imag], axis=-1))
    out = np.zeros((nz, ny, nx))
    rc = lib().ora_fft_c2r(nx, ny, nz, hin, out)
    assert rc == 0
    return out


def fft_line(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    n = x.shape[0]
    a = np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1))
    o = np.zeros_like(a)
    assert lib().ora_fft_c2c_line(n, a, o, int(inverse)) == 0
    return o[:, 0] + 1j * o[:, 1]


def twiddles(n: int) -> np.ndarray:
    w = np.zeros(2 * n)
    lib().ora_twiddles(n, w)
    return w[0::2] + 1j * w[1::2]


def kspace_table(nx: int, ny: int, nz: int):
    """kf, kf_sq, kf_4 in the reference's layout: arrays shaped (mx, ny, nz, 3) in
    C order == Fortran (3, nz, ny, mx)."""
    mx = nx // 2 + 1
    shp = (mx, ny, nz, 3, 2)
    a, b, c = (np.zeros(shp) for _ in range(3))
    lib().ora_kspace_table(nx, ny, nz, a, b, c)
    f = lambda t: t[..., 0] + 1j * t[..., 1]
    return f(a), f(b), f(c)


def gamma(var: int, s1: float, s2: float) -> np.ndarray:
    g = np.zeros(var * var)
    lib().ora_gamma(var, s1, s2, g)
    return g.reshape(var, var, order="F")


def make_params(**kw) -> Params:
    d = dict(t_step=0.01, grad_coeff_phi=0.5, grad_coeff_c=0.5, D_S=1.0, D_L=1.0, D_gb=0.0,
             mob_phi=1.0, sigma_1=1.5, sigma_2=1.5, am=0.5, ap=5.0, con_0_mat=0.1,
             con_0_ppt=1.0)
    d.update(kw)
    return Params(**d)


def run(params: Params, phi: np.ndarray, con: np.ndarray, nsteps: int):
    """Advance the state in place.  phi: (nz, ny, nx, var) (var fastest); con: (nz, ny, nx).
    Returns a structured list of per-step start-of-step diagnostics (dicts)."""
    nz, ny, nx, var = phi.shape
    assert phi.flags.c_contiguous and con.flags.c_contiguous
    diag = (Diag * max(nsteps, 1))()
    rc = lib().ora_run(C.byref(params), nx, ny, nz, var, phi, con, nsteps, C.cast(diag, C.c_void_p))
    assert rc == 0, rc
    return [{n: getattr(diag[i], n) for n in DIAG_FIELDS} for i in range(nsteps)]


def pointwise(params: Params, phi: np.ndarray, con: np.ndarray):
    """Loops 5-10 pointwise: phi (npts, var), con (npts,).  Returns
    (df_dc, df_dphi, mob_c, f_ch_sum, f_eta_sum)."""
    npts, var = phi.shape
    df_dc = np.zeros(npts)
    df_dphi = np.zeros((npts, var))
    mob = np.zeros(npts)
    fch, feta = C.c_double(), C.c_double()
    lib().ora_pointwise(C.byref(params), npts, var, np.ascontiguousarray(phi),
                        np.ascontiguousarray(con), df_dc, df_dphi, mob, C.byref(fch), C.byref(feta))
    return df_dc, df_dphi, mob, fch.value, feta.value


def extrema(phi: np.ndarray, con: np.ndarray):
    var = phi.shape[-1]
    out = np.zeros(4)
    lib().ora_extrema(con.size, var, np.ascontiguousarray(phi.reshape(-1, var)),
                      np.ascontiguousarray(con.reshape(-1)), out)
    return dict(zip(("c_max", "c_min", "phi_min", "phi_3_max"), out))


def ic_random(nx, ny, nz, var, iseed=-1, c0=0.5, c_amp=0.01, eta_amp=0.001):
    phi = np.zeros((nz, ny, nx, var))
    con = np.zeros((nz, ny, nx))
    lib().ora_ic_random(nx, ny, nz, var, iseed, c0, c_amp, eta_amp, phi, con)
    return phi, con


def ic_voronoi(nx, ny, nz, var, nseeds, iseed=-1, c0=0.4, c_amp=0.02):
    phi = np.zeros((nz, ny, nx, var))
    con = np.zeros((nz, ny, nx))
    lib().ora_ic_voronoi(nx, ny, nz, var, nseeds, iseed, c0, c_amp, phi, con)
    return phi, con


def set_threads(n: int) -> None:
    lib().ora_set_threads(int(n))
