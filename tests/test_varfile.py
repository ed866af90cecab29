"""trapbec_amd.varfile: the reference's data files (src/output.f90:48-107,
src/initial.f90:44,796-818) written and read back, gathered over ranks as bin/gather.m
does, and the driver's own output read through it."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from trapbec_amd import initial, varfile


def test_record_layout(tmp_path):
    """Fortran unformatted sequential records: 4-byte length, payload, 4-byte length."""
    p = tmp_path / "r.dat"
    varfile.write_records(p, [b"abc", b"defgh"])
    raw = p.read_bytes()
    assert raw == (np.int32(3).tobytes() + b"abc" + np.int32(3).tobytes() +
                   np.int32(5).tobytes() + b"defgh" + np.int32(5).tobytes())
    assert varfile.read_records(p) == [b"abc", b"defgh"]
    p.write_bytes(raw[:-1] + b"\x07")
    with pytest.raises(ValueError):
        varfile.read_records(p)


@pytest.mark.parametrize("prec", [64, 32])
def test_gather_four_ranks(tmp_path, prec):
    """A 2 x 2 decomposition written rank by rank (coords, L, t, block) gathers to the
    global array; restart files scatter into the two-digit directories and read back."""
    shape, dims = (8, 6, 4), (2, 2)
    cdt = np.complex128 if prec == 64 else np.complex64
    psi = (initial.noise(shape, 1, 1.0) + 0.5).astype(cdt)
    d = tmp_path / "data"
    for cx in range(2):
        for cy in range(2):
            r = cx * 2 + cy
            os.makedirs(d / f"proc{r:03d}")
            varfile.write_var(d / f"proc{r:03d}" / "var0003.dat",
                              psi[4 * cx:4 * cx + 4, 3 * cy:3 * cy + 3], (8.0, 6.0, 4.0),
                              0.25, (cx, cy), prec)
    varfile.write_dims(d, shape, dims)
    assert varfile.read_dims(d) == {"Nx": 8, "Ny": 6, "Nz": 4, "nprocx": 2, "nprocy": 2,
                                    "nmeshx": 4, "nmeshy": 3, "nmeshz": 4}
    g = varfile.gather(d, 3, prec)
    assert np.array_equal(g["psi"], psi) and g["t"] == 0.25 and g["L"] == (8.0, 6.0, 4.0)
    varfile.scatter_restart(d, psi, 41, 1.5, dims, prec)
    for r in range(4):
        cx, cy = divmod(r, 2)
        rs = varfile.read_restart(d / f"proc{r:02d}" / "var.dat", (4, 3, 4), prec)
        assert rs["itime"] == 41 and rs["t"] == 1.5
        assert np.array_equal(rs["psi"], psi[4 * cx:4 * cx + 4, 3 * cy:3 * cy + 3])


def test_gather_reference_four_rank_output(tmp_path):
    """varfile.gather on files the reference itself wrote: a 4-rank (2 x 2) run of
    12 x 10 x 8 FP64, its dims.log (src/output.f90:48-58) and every rank's var000N.dat
    (src/output.f90:71-78), as bin/gather.m assembles them.  Shot 1 equals the oracle's
    state after its 5 steps bit for bit (decomposition invariance), every shot carries
    the reference's t and L, and shots 2-4 differ from shot 1."""
    from conftest import GOLDEN, oracle
    f = np.load(os.path.join(GOLDEN, "info_f64_ragged_np4.npz"))
    d = tmp_path / "data"
    d.mkdir()
    (d / "dims.log").write_bytes(f["raw_dims_log"].tobytes())
    for key in f.files:
        if key.startswith("raw_proc"):
            rank, shot = key[4:11], key[12:]
            os.makedirs(d / rank, exist_ok=True)
            (d / rank / f"{shot}.dat").write_bytes(f[key].tobytes())
    assert varfile.read_dims(d) == {"Nx": 12, "Ny": 10, "Nz": 8, "nprocx": 2, "nprocy": 2,
                                    "nmeshx": 6, "nmeshy": 5, "nmeshz": 8}
    box = tuple(float(b) for b in f["box"])
    s = oracle.Stepper(oracle.params((12, 10, 8), box, tuple(float(h) for h in f["harm_w"]),
                                     dt=1e-3), f["psi0"])
    shots = int(f["shots"])
    prev = None
    for shot in range(1, int(f["steps"]) // shots + 1):
        g = varfile.gather(d, shot, 64)
        s.rk3(shots)
        assert g["L"] == box and g["t"] == pytest.approx(s.t, rel=1e-15)
        assert np.array_equal(g["psi"], s.interior), shot
        if prev is not None:
            assert not np.array_equal(g["psi"], prev)
        prev = g["psi"]


@pytest.mark.gpu
def test_driver_output_gathers(tmp_path):
    """The driver's dims.log and var0001.dat load through gather (one domain)."""
    drv = os.path.join(ROOT, "trapbec_amd", "driver", "bin", "tbec_run")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    nml = tmp_path / "run.nml"
    nml.write_text("&tbec_run\n Nx=16, Ny=12, Nz=8, Lx=16d0, Ly=12d0, Lz=8d0,\n"
                   " init_cond='single_vortex', dt_re=1d-3, nsteps=4, shots=4,\n"
                   f" out_dir='{tmp_path / 'data'}'\n/\n")
    print("\n[case] driver output gather", flush=True)
    r = subprocess.run([drv, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    g = varfile.gather(tmp_path / "data", 1)
    assert g["psi"].shape == (16, 12, 8) and g["t"] == pytest.approx(4e-3, rel=1e-12)
    rs = varfile.read_restart(tmp_path / "data" / "proc000" / "var.dat", (16, 12, 8))
    assert rs["itime"] == 5 and np.array_equal(rs["psi"], g["psi"])
