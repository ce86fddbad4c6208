#!/usr/bin/env python3
"""Writes the committed golden fixtures under tests/golden/.

  ran2_ref.json     ran_2 sequences from the reference's own src/ran_2.f90,
                    compiled with amdflang into oracle/_ref/libref.so
  kspace_ref.npz    kf, kf_sq, kf_4 of the reference's src/k_space.f90 (same
                    library) for a 12 x 8 x 10 grid, in its (3, Nz, Ny, Nx/2+1) order
  fft_numpy.npz     numpy.fft.rfftn vectors with f_trans normalisation (1/P)
  oracle_16.npz     a 16^3, var = 4 Voronoi state and the oracle's state and
                    summary rows after 20 FP64 steps (regression pin of the oracle)

Run in the build container (needs oracle/_ref, i.e. the reference tree):
  python3 tests/golden/make_golden.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402


def ran2_fixture():
    R = oracle.ref()
    assert R is not None, "oracle/_ref/libref.so is not built (make -C oracle ref)"
    out = {}
    for seed in (-1, -7, -123456):
        # the reference generator keeps SAVEd state; a non-positive idum re-seeds it
        idum = C.c_int(seed)
        vals = [R.ran_2_(C.byref(idum)) for _ in range(1000)]
        out[str(seed)] = dict(values=vals, idum_after=idum.value)
    json.dump(out, open(os.path.join(HERE, "ran2_ref.json"), "w"))


def kspace_fixture():
    R = oracle.ref()
    nx, ny, nz = 12, 8, 10
    mx = nx // 2 + 1
    n = 3 * nz * ny * mx
    bufs = [(C.c_double * (2 * n))() for _ in range(3)]
    fst = (C.c_int * 3)(1, 1, 1)
    fen = (C.c_int * 3)(nz, ny, mx)
    R.k_space_(C.byref(C.c_int(nx)), C.byref(C.c_int(ny)), C.byref(C.c_int(nz)), fst, fen, *bufs)
    arr = [np.frombuffer(b, dtype=np.float64).copy() for b in bufs]
    np.savez(os.path.join(HERE, "kspace_ref.npz"), nx=nx, ny=ny, nz=nz, kf=arr[0], kf_sq=arr[1], kf_4=arr[2])


def fft_fixture():
    rng = np.random.default_rng(2024)
    d = {}
    for shape in ((8, 12, 16), (10, 6, 8)):  # (nz, ny, nx)
        x = rng.standard_normal(shape)
        d[f"x_{shape[2]}_{shape[1]}_{shape[0]}"] = x
        d[f"X_{shape[2]}_{shape[1]}_{shape[0]}"] = np.fft.rfftn(x) / x.size
    np.savez(os.path.join(HERE, "fft_numpy.npz"), **d)


def oracle_fixture():
    phi, con = oracle.ic_voronoi(16, 16, 16, 4, 6)
    p = oracle.make_params(t_step=0.005, grad_coeff_phi=1.0, grad_coeff_c=1.0, D_gb=0.2)
    phi0, con0 = phi.copy(), con.copy()
    diag = oracle.run(p, phi, con, 20)
    rows = np.array([[d[k] for k in oracle.DIAG_FIELDS] for d in diag])
    np.savez_compressed(os.path.join(HERE, "oracle_16.npz"), phi0=phi0, con0=con0, phi=phi, con=con,
                        rows=rows)


if __name__ == "__main__":
    ran2_fixture()
    kspace_fixture()
    fft_fixture()
    oracle_fixture()
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))
