"""Restatements of the reference's initial conditions, precipitate insertion and
checkpoint format, for checking the host driver (test infrastructure only).

  circle            src/var_diff.f90:357-449
  plane             src/var_diff.f90:254-355
  insert            src/var_diff.f90:479-554
  gamma             src/var_diff.f90:206-218
  read/write_fortran_record   Fortran unformatted sequential files with 4-byte
                    record markers (the layout of data_save.NNNN and
                    data_rerun.NNNN, src/var_diff.f90:453-457, 1134-1138)

Arrays use C order: phi (nz, ny, nx, var), con (nz, ny, nx).
"""
from __future__ import annotations

import struct

import numpy as np

from . import Ran2


def _f32_sq_ratio(i, r):
    """(float(i)/float(r))**2 evaluated in single precision, then widened."""
    return float(np.float32(np.float32(i) / np.float32(r)) ** np.float32(2))


def _ellipsoid_offsets(rad):
    out = []
    for k in range(-rad[2], rad[2] + 1):
        for j in range(-rad[1], rad[1] + 1):
            for i in range(-rad[0], rad[0] + 1):
                s = _f32_sq_ratio(i, rad[0]) + _f32_sq_ratio(j, rad[1]) + _f32_sq_ratio(k, rad[2])
                if s <= 1.0:
                    out.append((i, j, k))
    return out


def _mark(phi, centre, offs):
    nz, ny, nx, _ = phi.shape
    ci, cj, ck = centre
    for (i, j, k) in offs:
        ii, jj, kk = i + ci, j + cj, k + ck  # 1-based
        if 1 <= ii <= nx and 1 <= jj <= ny and 1 <= kk <= nz:
            phi[kk - 1, jj - 1, ii - 1, 2] = 1.0
            phi[kk - 1, jj - 1, ii - 1, 0] = 0.0
            phi[kk - 1, jj - 1, ii - 1, 1] = 0.0


def plane(cfg, phi):
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    j_ppt = cfg["j_ppt"]
    j1 = ny - j_ppt + 1
    phi[..., 1] = 1.0
    y = np.arange(1, ny + 1)
    for mask in (y < j1, y >= j_ppt):
        phi[:, mask, :, :] = 0.0
        phi[:, mask, :, 0] = 1.0
    offs = _ellipsoid_offsets(cfg["ppt_rad"])
    _mark(phi, (cfg["i_ppt"], j_ppt, cfg["k_ppt"]), offs)
    _mark(phi, (cfg["i_ppt"], ny - j_ppt + 1, cfg["k_ppt"]), offs)


def _dist2(nx, ny, nz, ix, jy, kz):
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    return (i - ix) ** 2 + (j - jy) ** 2 + (k - kz) ** 2


def _draw(rng, n):
    # ix = ran_2(iseed)*Nx + 1.0, truncated to integer
    return int(rng() * n + 1.0)


def circle(cfg, phi, con, rng):
    nx, ny, nz, var = cfg["nx"], cfg["ny"], cfg["nz"], cfg["var"]
    g = cfg["grn_rad"]
    phi[..., 0] = 1.0
    for (i, j, k) in _ellipsoid_offsets((g, g, g)):
        ii, jj, kk = i + nx // 2, j + ny // 2, k + nz // 2
        if 1 <= ii <= nx and 1 <= jj <= ny and 1 <= kk <= nz:
            phi[kk - 1, jj - 1, ii - 1, 1] = 1.0
            phi[kk - 1, jj - 1, ii - 1, 0] = 0.0
            phi[kk - 1, jj - 1, ii - 1, 2] = 0.0
    rad = sum(r * r for r in cfg["ppt_rad"])
    nppt = 1
    while nppt <= cfg["num_ppt"]:
        ix, jy, kz = _draw(rng, nx), _draw(rng, ny), _draw(rng, nz)
        inside = _dist2(nx, ny, nz, ix, jy, kz) <= rad
        if np.any(inside & (phi[..., var - 1] > 0.0)):
            continue
        phi[inside, 2] = 1.0
        phi[inside, 1] = 0.0
        phi[inside, 0] = 0.0
        con[inside] = cfg["con_0_ppt"]
        nppt += 1


def insert(cfg, phi, rng, max_draws=1_000_000):
    """The reference redraws until a centre passes both tests and has no bound
    on the number of draws (it never ends when no site qualifies, e.g. once the
    precipitate field is non-zero everywhere); here max_draws ends it with an error."""
    nx, ny, nz, var = cfg["nx"], cfg["ny"], cfg["nz"], cfg["var"]
    rad = sum(r * r for r in cfg["ppt_rad"])
    nppt = 1
    draws = 0
    while nppt <= cfg["num_ppt"]:
        while True:
            draws += 1
            if draws > max_draws:
                raise RuntimeError("no admissible precipitate site found")
            ix, jy, kz = _draw(rng, nx), _draw(rng, ny), _draw(rng, nz)
            s = 0.0
            if 1 <= ix <= nx and 1 <= jy <= ny and 1 <= kz <= nz:
                for v in range(var - 1):
                    fi = phi[kz - 1, jy - 1, ix - 1, v]
                    s = s + fi * fi
            if s > float(np.float32(0.8)):
                continue
            inside = _dist2(nx, ny, nz, ix, jy, kz) <= rad
            if np.any(inside & (phi[..., var - 1] > 0.0)):
                continue
            break
        phi[inside, var - 1] = 1.0
        phi[inside, : var - 1] = 0.0
        nppt += 1


def gamma(var, s1, s2):
    g = np.zeros((var, var))
    for j in range(var):
        for i in range(var):
            if i != j and i < var - 1 and j < var - 1:
                g[i, j] = s1
            if (j == var - 1 and i < var - 1) or (i == var - 1 and j < var - 1):
                g[i, j] = s2
    return g


# ---------------------------------------------------------------------------
def read_fortran_records(path):
    """Records of a Fortran unformatted sequential file: 4-byte length markers
    around each (sub)record; the leading marker is negative when the record
    continues in a further subrecord, the trailing marker negative when the
    subrecord continues a previous one."""
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    parts = []
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        ln = abs(n)
        assert pos + 8 + ln <= len(data), "truncated record"
        (tail,) = struct.unpack_from("<i", data, pos + 4 + ln)
        assert abs(tail) == ln, "corrupt record markers"
        assert (tail < 0) == (len(parts) > 0), "trailing marker sign does not match continuation"
        parts.append(data[pos + 4: pos + 4 + ln])
        pos += 8 + ln
        if n >= 0:
            recs.append(parts[0] if len(parts) == 1 else b"".join(parts))
            parts = []
    assert not parts, "truncated continued record"
    return recs


MAX_SUBRECORD = 2147483639


def write_fortran_records(path, recs):
    """Inverse of read_fortran_records; records longer than MAX_SUBRECORD bytes
    are split into subrecords (leading marker negative when more follow,
    trailing marker negative when one precedes)."""
    with open(path, "wb") as f:
        for r in recs:
            mv = memoryview(r)
            done = 0
            while True:
                sub = min(MAX_SUBRECORD, len(mv) - done)
                more = done + sub < len(mv)
                f.write(struct.pack("<i", -sub if more else sub))
                f.write(mv[done:done + sub])
                f.write(struct.pack("<i", -sub if done > 0 else sub))
                done += sub
                if not more:
                    break


def read_checkpoint(path, nx, ny, nz, var):
    """data_save.NNNN -> (ist, ien, phi (nz,ny,nx,var), con (nz,ny,nx))."""
    r = read_fortran_records(path)
    assert len(r) == 2, f"{path}: expected 2 records, found {len(r)}"
    b = np.frombuffer(r[0], dtype="<i4")
    body = np.frombuffer(r[1], dtype="<f8")
    P = nx * ny * nz
    assert body.size == P * (var + 1), "record 2 has the wrong length"
    phi = body[: P * var].reshape(nz, ny, nx, var).copy()
    con = body[P * var:].reshape(nz, ny, nx).copy()
    return b[:3].copy(), b[3:].copy(), phi, con


def write_restart(path, phi, con):
    nz, ny, nx, var = phi.shape
    r1 = np.array([1, 1, 1, nx, ny, nz], dtype="<i4").tobytes()
    r2 = np.ascontiguousarray(phi, dtype="<f8").tobytes() + np.ascontiguousarray(con, dtype="<f8").tobytes()
    write_fortran_records(path, [r1, r2])


def analyze(phi, threshold):
    """The per-voxel quantities and counts of src/paraview.f90:188-313 for one
    state phi (nz, ny, nx, var); arrays returned flat, x fastest.  The gradient
    faces (left unset by the reference) are 0 here."""
    nz, ny, nx, var = phi.shape
    sq = (phi * phi).sum(axis=-1)
    above = (phi > threshold).sum(axis=-1)
    big = phi > 0.8
    vmax = np.where(big.any(axis=-1), var - np.argmax(big[..., ::-1], axis=-1), 0)
    bnd = above > 1
    p3 = phi[..., 2] if var >= 3 else np.zeros((nz, ny, nx))
    g = np.zeros((nz, ny, nx))
    c = p3
    gx = (c[1:-1, 1:-1, 2:] - c[1:-1, 1:-1, :-2]) / 2.0
    gy = (c[1:-1, 2:, 1:-1] - c[1:-1, :-2, 1:-1]) / 2.0
    gz = (c[2:, 1:-1, 1:-1] - c[:-2, 1:-1, 1:-1]) / 2.0
    g[1:-1, 1:-1, 1:-1] = np.sqrt(gx * gx + gy * gy + gz * gz)
    counts = (int((bnd & (phi[..., var - 1] > threshold)).sum()), int(bnd.sum()))
    return dict(phi_sq=sq.ravel(), var_num=vmax.ravel(), phi_3=p3.ravel(), grad_phi_3=g.ravel(),
                boundary=bnd.astype(float).ravel(), counts=counts)


def read_vtk_values(path):
    """Values of an ASCII STRUCTURED_POINTS file written by the post-processor."""
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[9] == "LOOKUP_TABLE default"
    dims = [int(x) for x in lines[4].split()[1:]]
    vals = np.array(" ".join(lines[10:]).split(), dtype=float)
    assert vals.size == dims[0] * dims[1] * dims[2]
    return dims, vals
