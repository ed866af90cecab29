"""The Fortran host driver (trapbec_amd/driver): namelist, seeded initial state, var and
restart files in the reference's record layout (src/output.f90:71-108), restart
(src/initial.f90:796-818) and errors in the emergency_stop style."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, oracle
from trapbec_amd import initial

DRV = os.path.join(ROOT, "trapbec_amd", "driver", "bin", "tbec_run")


def _need_driver():
    if not os.path.exists(DRV):
        pytest.skip("driver not built")


def _nml(path, **kw):
    lines = ["&tbec_run"]
    for k, v in kw.items():
        if isinstance(v, bool):
            v = ".true." if v else ".false."
        elif isinstance(v, str):
            v = f"'{v}'"
        elif isinstance(v, float):
            v = repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
        lines.append(f"  {k}={v},")
    lines.append("/")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _read_records(path):
    out = []
    data = open(path, "rb").read()
    pos = 0
    while pos < len(data):
        n = int(np.frombuffer(data[pos:pos + 4], np.int32)[0])
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def _read_var(path, shape, prec=64):
    rdt = np.float64 if prec == 64 else np.float32
    cdt = np.complex128 if prec == 64 else np.complex64
    rec = _read_records(path)[0]
    coords = np.frombuffer(rec[:8], np.int32)
    hdr = np.frombuffer(rec[8:8 + 4 * rdt().itemsize], rdt)
    psi = np.frombuffer(rec[8 + 4 * rdt().itemsize:], cdt).reshape(shape)
    return coords, hdr, psi


C1_16 = dict(Nx=16, Ny=16, Nz=16, Lx=16.0, Ly=16.0, Lz=16.0, init_cond="tf_vortex",
             mu=10.0, g_phys=500.0, n_lines=1, noise_amp=1e-3, seed=20261203,
             psi_scale=math.sqrt(1000.0))


def test_driver_initial_state_matches_python(tmp_path):
    """The Fortran seeded initial state equals trapbec_amd.initial's (counter RNG, TF,
    tanh vortex, noise) to rounding."""
    _need_driver()
    nml = tmp_path / "a.nml"
    _nml(nml, **C1_16, ic_only=True, out_dir=str(tmp_path / "data"))
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    coords, hdr, psi = _read_var(tmp_path / "data" / "proc000" / "var0000.dat", (16, 16, 16))
    assert list(coords) == [0, 0] and list(hdr[:3]) == [16.0, 16.0, 16.0]
    _, ref = initial.config_c1(n=16)
    assert np.abs(psi - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("dt, warned", [(0.01, False), (0.2, True)])
def test_driver_timestep_warning(tmp_path, dt, warned):
    """The reference's stability check (src/initial.f90:29-33): dt above min(h)^2 / 8
    warns, else the limit is printed (16^3 on a box of 16: h = 1, limit 0.125)."""
    _need_driver()
    nml = tmp_path / "dt.nml"
    _nml(nml, **dict(C1_16, dt_re=dt), ic_only=True, out_dir=str(tmp_path / "data"))
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ("WARNING TIMESTEP IS PROBABLY TOO LARGE!" in r.stdout) == warned
    if not warned:
        line = next(ln for ln in r.stdout.splitlines() if "below upper-limit" in ln)
        assert float(line.split()[-1]) == pytest.approx(0.125)


def test_driver_errors_like_emergency_stop(tmp_path):
    _need_driver()
    nml = tmp_path / "bad.nml"
    _nml(nml, **dict(C1_16, init_cond="no_such_state"), ic_only=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0
    assert "FATAL ERROR" in r.stdout
    assert "wrong initial conditions set" in (tmp_path / "ERROR").read_text()


@pytest.mark.gpu
def test_driver_run_matches_oracle(tmp_path):
    """20 real-time steps through the driver equal the oracle (FP64, 1e-12)."""
    _need_driver()
    nml = tmp_path / "run.nml"
    _nml(nml, **C1_16, nsteps=20, shots=20, out_dir=str(tmp_path / "data"))
    print(f"\n[case] driver run 16^3", flush=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, hdr, psi = _read_var(tmp_path / "data" / "proc000" / "var0001.dat", (16, 16, 16))
    pr, psi0 = initial.config_c1(n=16)
    s = oracle.Stepper(oracle.params(pr.shape, pr.box, pr.kernel_params()["harm_w"],
                                     dt=1e-3), psi0)
    s.rk3(20)
    assert np.abs(psi - s.interior).max() < 1e-12 * np.abs(s.interior).max()
    assert hdr[3] == pytest.approx(0.02, rel=1e-14)


@pytest.mark.gpu
def test_driver_reference_imaginary_stage(tmp_path):
    """The reference's imaginary stage through the driver (imag_scheme='euler',
    src/run.f90:19-33): 20 unnormalised Euler steps from the reference's own soliton state,
    shots = 10.  The imag_var0002.dat snapshot (src/output.f90:62-69) equals the
    reference's own imag_var0001.dat of a 20-step, shots = 20 run (the euler_f64_16
    fixture) to 1e-12, and the printed norm ratios intPsi2/intPsi2_old of steps 10 and 20
    equal the oracle's to 1e-11."""
    _need_driver()
    f = np.load(os.path.join(GOLDEN, "euler_f64_16.npz"))
    box = tuple(float(b) for b in f["box"])
    harm = tuple(float(h) for h in f["harm_w"])
    nml = tmp_path / "imag.nml"
    _nml(nml, Nx=16, Ny=16, Nz=16, Lx=box[0], Ly=box[1], Lz=box[2], harm_wx=harm[0],
         harm_wy=harm[1], harm_wz=harm[2], dt_re=float(f["dt"]), init_cond="soliton",
         imaginary_time=True, imag_scheme="euler", imag_nsteps=20, shots=10,
         imaginary_output=True, nsteps=0, out_dir=str(tmp_path / "data"))
    print("\n[case] driver reference imaginary stage 16^3", flush=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, hdr, psi = _read_var(tmp_path / "data" / "proc000" / "imag_var0002.dat", (16, 16, 16))
    ref = f["psi"]
    assert np.abs(psi - ref).max() <= 1e-12 * np.abs(ref).max()
    assert hdr[3] == 0.0                       # imaginary time does not advance t
    rows = [ln.split() for ln in r.stdout.splitlines()]
    ratios = [float(t[1]) for t in rows if len(t) == 2 and t[0] in ("1", "2")]
    psi0 = initial.reference_soliton((16, 16, 16), box, harm)
    s = oracle.Stepper(oracle.params((16, 16, 16), box, harm, dt=float(f["dt"])), psi0)
    want = []
    for _ in range(2):
        s.euler(9)
        n_old = float(np.sum(np.abs(s.interior) ** 2))
        s.euler(1)
        want.append(float(np.sum(np.abs(s.interior) ** 2)) / n_old)
    assert len(ratios) == 2
    assert ratios == pytest.approx(want, rel=1e-11)


@pytest.mark.gpu
def test_driver_restart_is_bitwise(tmp_path):
    """Stop after 10 steps, restart from var.dat, run to 20: identical to 20 straight."""
    _need_driver()
    print(f"\n[case] driver restart 16^3", flush=True)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    _nml(a / "run.nml", **C1_16, nsteps=20, shots=10, out_dir=str(a / "data"))
    assert subprocess.run([DRV, "run.nml"], cwd=a, capture_output=True).returncode == 0
    _nml(b / "run.nml", **C1_16, nsteps=10, shots=10, out_dir=str(b / "data"))
    assert subprocess.run([DRV, "run.nml"], cwd=b, capture_output=True).returncode == 0
    _nml(b / "run2.nml", **dict(C1_16, init_cond="restart"), nsteps=20, shots=10,
         out_dir=str(b / "data"))
    r = subprocess.run([DRV, "run2.nml"], cwd=b, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    _, ha, pa = _read_var(a / "data" / "proc000" / "var0002.dat", (16, 16, 16))
    _, hb, pb = _read_var(b / "data" / "proc000" / "var0002.dat", (16, 16, 16))
    assert np.array_equal(pa, pb)
    assert ha[3] == hb[3]


@pytest.mark.gpu
def test_driver_writes_velocity_and_qpressure(tmp_path):
    """vel_print / qpress_print write vel01.dat and qpress01.dat (src/output.f90:80-91:
    coords, field(nz, ny, nx, 3)) from the state of var0001.dat; they equal the oracle's
    restatement of get_phase / get_qpressure on that state."""
    _need_driver()
    nml = tmp_path / "run.nml"
    _nml(nml, **C1_16, nsteps=5, shots=5, vel_print=True, qpress_print=True,
         out_dir=str(tmp_path / "data"))
    print(f"\n[case] driver vel/qpress 16^3", flush=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    d = tmp_path / "data" / "proc000"
    _, _, psi = _read_var(d / "var0001.dat", (16, 16, 16))
    s = oracle.Stepper(oracle.params((16, 16, 16), 16.0, (256.0,) * 3), psi)
    for name, ref in (("vel", s.velocity()), ("qpress", s.qpressure())):
        rec = _read_records(d / f"{name}01.dat")[0]
        assert list(np.frombuffer(rec[:8], np.int32)) == [0, 0]
        got = np.frombuffer(rec[8:], np.float64).reshape(3, 16, 16, 16)
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), name


@pytest.mark.gpu
def test_driver_solid_boundaries_match_reference(tmp_path):
    """boundaries='solid' through the driver's namelist: 20 steps from the solid-boundary
    reference fixture's psi0 equal the reference's output (FP64, 1e-12)."""
    _need_driver()
    f = np.load(os.path.join(ROOT, "tests", "golden", "rk3_f64_16_solid.npz"))
    d = tmp_path / "data" / "proc000"
    d.mkdir(parents=True)
    # restart file: [itime] [t] [Psi] (src/initial.f90:796-818)
    with open(d / "var.dat", "wb") as fh:
        for payload in (np.array([1], np.int32).tobytes(), np.array([0.0]).tobytes(),
                        np.ascontiguousarray(f["psi0"]).tobytes()):
            n = np.array([len(payload)], np.int32).tobytes()
            fh.write(n + payload + n)
    box = tuple(float(b) for b in f["box"])
    h = tuple(float(x) for x in f["harm_w"])
    nml = tmp_path / "run.nml"
    _nml(nml, Nx=16, Ny=16, Nz=16, Lx=box[0], Ly=box[1], Lz=box[2], harm_wx=h[0],
         harm_wy=h[1], harm_wz=h[2], omega_z=float(f["omega_z"]), init_cond="restart",
         boundaries="solid", nsteps=20, shots=20, out_dir=str(tmp_path / "data"))
    print(f"\n[case] driver solid boundaries 16^3", flush=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, psi = _read_var(d / "var0001.dat", (16, 16, 16))
    ref = f["psi"]
    assert np.abs(psi - ref).max() < 1e-12 * np.abs(ref).max()


def _write_restart(d, psi0, t0, itime=1):
    """var.dat in the reference's restart layout: [itime] [t] [Psi] (src/initial.f90:796-818)."""
    d.mkdir(parents=True, exist_ok=True)
    rdt = np.float64 if psi0.dtype == np.complex128 else np.float32
    with open(d / "var.dat", "wb") as fh:
        for payload in (np.array([itime], np.int32).tobytes(), np.array([t0], rdt).tobytes(),
                        np.ascontiguousarray(psi0).tobytes()):
            n = np.array([len(payload)], np.int32).tobytes()
            fh.write(n + payload + n)


def test_driver_rejects_unknown_potential(tmp_path):
    _need_driver()
    nml = tmp_path / "bad.nml"
    _nml(nml, **dict(C1_16, potential="no_such_trap"))
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode != 0
    assert "wrong potential set" in (tmp_path / "ERROR").read_text()


@pytest.mark.gpu
def test_driver_box_trap_matches_reference(tmp_path):
    """potential='box_trap' with the frame term spun down mid-run (t_spindown), restarted
    at t = 15 through the namelist: the reference fixture's output (FP64, 1e-12)."""
    _need_driver()
    f = np.load(os.path.join(ROOT, "tests", "golden", "pot_box_f64.npz"))
    shape = f["psi0"].shape
    _write_restart(tmp_path / "data" / "proc000", f["psi0"], float(f["t0"]))
    box = tuple(float(b) for b in f["box"])
    h = tuple(float(x) for x in f["harm_w"])
    nml = tmp_path / "run.nml"
    _nml(nml, Nx=shape[0], Ny=shape[1], Nz=shape[2], Lx=box[0], Ly=box[1], Lz=box[2],
         harm_wx=h[0], harm_wy=h[1], harm_wz=h[2], omega_z=float(f["omega_z"]),
         init_cond="restart", potential="box_trap", t_spindown=float(f["t_spindown"]),
         nsteps=20, shots=20, out_dir=str(tmp_path / "data"))
    print(f"\n[case] driver box_trap {shape}", flush=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, hdr, psi = _read_var(tmp_path / "data" / "proc000" / "var0001.dat", shape)
    ref = f["psi"]
    assert np.abs(psi - ref).max() < 1e-12 * np.abs(ref).max()
    assert hdr[3] == pytest.approx(float(f["t"]), rel=1e-15)


@pytest.mark.gpu
def test_driver_forcing_matches_oracle(tmp_path):
    """forcing with force_freq = 5 (src/run.f90:70-74): a vortex line is imprinted after
    steps 5 and 10 at seeded positions in [-L/4, L/4]^2; the step-10 snapshot is written
    before that step's imprint, as in the reference's loop.  Against the oracle stepping
    and imprinting at the same draws (FP64, 1e-12)."""
    _need_driver()
    nml = tmp_path / "run.nml"
    _nml(nml, **C1_16, nsteps=15, shots=10, forcing=True, force_freq=5,
         out_dir=str(tmp_path / "data"))
    print("\n[case] driver forcing 16^3", flush=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, psi = _read_var(tmp_path / "data" / "proc000" / "var0001.dat", (16, 16, 16))
    pr, psi0 = initial.config_c1(n=16)
    s = oracle.Stepper(oracle.params(pr.shape, pr.box, pr.kernel_params()["harm_w"],
                                     dt=1e-3), psi0)
    seed = C1_16["seed"]
    s.rk3(5)
    u = initial.uniform(seed, 7, np.arange(3))
    s.imprint_vortex(-4.0 + u[0] * 8.0, -4.0 + u[1] * 8.0, -1.0 if u[2] < 0.5 else 1.0)
    s.rk3(5)
    ref = s.interior
    assert np.abs(psi - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ic_crow_f64", "ins_dipole_f64", "ins_pcurrent_barrier_f64"])
def test_driver_reference_states_match_reference(tmp_path, name):
    """init_cond / add_pcurrent / add_dipole through the namelist, one step with dt = 0:
    var0001.dat equals the reference's snapshot (FP64, 1e-12)."""
    _need_driver()
    f = np.load(os.path.join(ROOT, "tests", "golden", name + ".npz"))
    shape = f["psi"].shape
    box = tuple(float(b) for b in f["box"])
    h = tuple(float(x) for x in f["harm_w"])
    nml = tmp_path / "run.nml"
    _nml(nml, Nx=shape[0], Ny=shape[1], Nz=shape[2], Lx=box[0], Ly=box[1], Lz=box[2],
         harm_wx=h[0], harm_wy=h[1], harm_wz=h[2], init_cond=str(f["init_cond"]),
         potential=str(f["potential"]), add_dipole=bool(f["add_dipole"]),
         add_pcurrent=bool(f["add_pcurrent"]), dt_re=0.0, nsteps=1, shots=1,
         out_dir=str(tmp_path / "data"))
    print(f"\n[case] driver {name}", flush=True)
    r = subprocess.run([DRV, str(nml)], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, _, psi = _read_var(tmp_path / "data" / "proc000" / "var0001.dat", shape)
    ref = f["psi"]
    assert np.abs(psi - ref).max() < 1e-12 * np.abs(ref).max()
