"""The reference's data files, read and written from Python: per-rank snapshots
``data/procNNN/var%04d.dat`` (one record: coords(2) int32, Lx, Ly, Lz, t, Psi;
src/output.f90:71-78), restart files ``var.dat`` (records [itime] [t] [Psi];
src/output.f90:103-107, read by reload_data from ``data/procNN/var.dat``,
src/initial.f90:44,796-818), the velocity / quantum-pressure records
(src/output.f90:80-91) and ``dims.log`` (src/output.f90:48-58).

gather() assembles a P-rank run into one array, as bin/gather.m does, so results of the
reference's MPI runs load into a single-GPU handle; scatter_restart() writes the
per-rank restart files a P-rank reference run reads, in the two-digit directories
reload_data opens.  Arrays are numpy, indexed [x, y, z] (C order), the memory order of
the reference's Psi(k, j, i).  Records are Fortran unformatted sequential records with
4-byte length markers.
"""
from __future__ import annotations

import os

import numpy as np


def _rdt(prec):
    return np.float64 if prec == 64 else np.float32


def _cdt(prec):
    return np.complex128 if prec == 64 else np.complex64


def read_records(path: str) -> list:
    data = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(data):
        n = int(np.frombuffer(data[pos:pos + 4], np.int32)[0])
        tail = int(np.frombuffer(data[pos + 4 + n:pos + 8 + n], np.int32)[0])
        if tail != n:
            raise ValueError(f"{path}: record markers {n} / {tail} disagree")
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def write_records(path: str, payloads) -> None:
    with open(path, "wb") as f:
        for p in payloads:
            n = np.array([len(p)], np.int32).tobytes()
            f.write(n + p + n)


def read_dims(data_dir: str) -> dict:
    """dims.log: Nx, Ny, Nz, nprocx, nprocy, nmeshx, nmeshy, nmeshz (one per line)."""
    v = [int(float(x)) for x in open(os.path.join(data_dir, "dims.log")).read().split()]
    keys = ("Nx", "Ny", "Nz", "nprocx", "nprocy", "nmeshx", "nmeshy", "nmeshz")
    return dict(zip(keys, v))


def write_dims(data_dir: str, shape, dims=(1, 1)) -> None:
    nx, ny, nz = shape
    vals = [nx, ny, nz, dims[0], dims[1], nx // dims[0], ny // dims[1], nz]
    with open(os.path.join(data_dir, "dims.log"), "w") as f:
        f.write("".join(f"{v:12d}\n" for v in vals))


def read_var(path: str, block_shape, prec: int = 64) -> dict:
    """One snapshot record: coords, (Lx, Ly, Lz), t and the rank's block of psi."""
    rdt = _rdt(prec)
    rec = read_records(path)[0]
    coords = np.frombuffer(rec[:8], np.int32)
    hdr = np.frombuffer(rec[8:8 + 4 * rdt().itemsize], rdt)
    psi = np.frombuffer(rec[8 + 4 * rdt().itemsize:], _cdt(prec)).reshape(block_shape)
    return {"coords": (int(coords[0]), int(coords[1])), "L": tuple(float(x) for x in hdr[:3]),
            "t": float(hdr[3]), "psi": psi}


def write_var(path: str, psi, L, t, coords=(0, 0), prec: int = 64) -> None:
    rdt = _rdt(prec)
    payload = (np.array(coords, np.int32).tobytes() +
               np.array([L[0], L[1], L[2], t], rdt).tobytes() +
               np.ascontiguousarray(psi, _cdt(prec)).tobytes())
    write_records(path, [payload])


def gather(data_dir: str, shot: int, prec: int = 64, prefix: str = "var") -> dict:
    """Assemble snapshot `shot` of every rank (bin/gather.m): psi [x, y, z], t, L."""
    d = read_dims(data_dir)
    bx, by, bz = d["nmeshx"], d["nmeshy"], d["nmeshz"]
    psi = np.zeros((d["Nx"], d["Ny"], d["Nz"]), _cdt(prec))
    ts, L = [], None
    for r in range(d["nprocx"] * d["nprocy"]):
        v = read_var(os.path.join(data_dir, f"proc{r:03d}", f"{prefix}{shot:04d}.dat"),
                     (bx, by, bz), prec)
        cx, cy = v["coords"]
        psi[cx * bx:(cx + 1) * bx, cy * by:(cy + 1) * by, :] = v["psi"]
        ts.append(v["t"])
        L = v["L"]
    if len(set(ts)) != 1:
        raise ValueError(f"t not consistent across ranks: {ts}")
    return {"psi": psi, "t": ts[0], "L": L}


def gather_field(data_dir: str, name: str, shot: int, prec: int = 64) -> np.ndarray:
    """vel%02d.dat / qpress%02d.dat of every rank (coords, field(nz, ny, nx, 3)) as an
    array indexed [component, x, y, z]."""
    d = read_dims(data_dir)
    bx, by, bz = d["nmeshx"], d["nmeshy"], d["nmeshz"]
    out = np.zeros((3, d["Nx"], d["Ny"], d["Nz"]), _rdt(prec))
    for r in range(d["nprocx"] * d["nprocy"]):
        rec = read_records(os.path.join(data_dir, f"proc{r:03d}", f"{name}{shot:02d}.dat"))[0]
        cx, cy = (int(c) for c in np.frombuffer(rec[:8], np.int32))
        blk = np.frombuffer(rec[8:], _rdt(prec)).reshape(3, bx, by, bz)
        out[:, cx * bx:(cx + 1) * bx, cy * by:(cy + 1) * by, :] = blk
    return out


def read_restart(path: str, shape, prec: int = 64) -> dict:
    """var.dat: itime, t, psi (src/initial.f90:796-818)."""
    recs = read_records(path)
    return {"itime": int(np.frombuffer(recs[0], np.int32)[0]),
            "t": float(np.frombuffer(recs[1], _rdt(prec))[0]),
            "psi": np.frombuffer(recs[2], _cdt(prec)).reshape(shape)}


def write_restart(path: str, itime: int, t: float, psi, prec: int = 64) -> None:
    write_records(path, [np.array([itime], np.int32).tobytes(),
                         np.array([t], _rdt(prec)).tobytes(),
                         np.ascontiguousarray(psi, _cdt(prec)).tobytes()])


def scatter_restart(data_dir: str, psi, itime: int, t: float, dims=(1, 1),
                    prec: int = 64) -> None:
    """Per-rank restart files for a reference run on dims[0] x dims[1] ranks: rank
    r = cx*dims[1] + cy holds block (cx, cy), in data/procNN/var.dat (the two-digit
    directory reload_data reads)."""
    nx, ny, _ = psi.shape
    bx, by = nx // dims[0], ny // dims[1]
    if bx * dims[0] != nx or by * dims[1] != ny:
        raise ValueError("grid does not divide over the ranks")
    for cx in range(dims[0]):
        for cy in range(dims[1]):
            r = cx * dims[1] + cy
            d = os.path.join(data_dir, f"proc{r:02d}")
            os.makedirs(d, exist_ok=True)
            write_restart(os.path.join(d, "var.dat"), itime, t,
                          psi[cx * bx:(cx + 1) * bx, cy * by:(cy + 1) * by, :], prec)
