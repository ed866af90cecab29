"""Write the reference golden fixtures under tests/golden/ (run in the build container).

Each case compiles the reference Fortran (oracle/refbuild.py: symlinks to
/root/reference/src, a generated cparam.local, amdflang -O3 -ffp-contract=off, MPICH),
injects a seeded psi0 through the restart file, runs it, and stores psi0 (or its
generator parameters) and the reference output.  Only these data files are committed;
the reference itself never travels.

    python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)
import refbuild  # noqa: E402
from trapbec_amd import initial  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def small_ic(shape, box, seed, scale=True):
    """TF (mu = 10, V = r^2/2 in physical units, g = 500) times a z-line vortex plus
    1e-3 noise, in reference units (psi * sqrt(2g))."""
    psi, _ = initial.tf_profile(shape, box, (1.0, 1.0, 1.0), 500.0, mu=10.0)
    psi = psi * initial.vortex_line_factor(shape, box, 0.3, -0.2, 1, 1 / math.sqrt(10))
    psi = psi + initial.noise(shape, seed, 1e-3)
    return psi * math.sqrt(1000.0) if scale else psi


def case_small(name, prec, shape=(16, 16, 16), box=(16.0, 16.0, 16.0), steps=20,
               omega_z=0.0, gamma=0.0, dt=1e-3 + 0j, nprocs=1, seed=11, boundaries="zero",
               periodicz=False, offset=(0.0, 0.0)):
    harm = tuple(L * L for L in box)   # pot = x^2 + y^2 + z^2, i.e. 2V with w = 1
    d = refbuild.build(name, prec=prec, Nx=shape[0], Ny=shape[1], Nz=shape[2],
                       Lx=box[0], Ly=box[1], Lz=box[2], nsteps=steps, shots=steps,
                       omega_z=omega_z, gamma=gamma, dt_re=dt.real, dt_im=dt.imag,
                       harm_wx=harm[0], harm_wy=harm[1], harm_wz=harm[2],
                       boundaries=boundaries, periodicz=periodicz)
    psi0 = small_ic(shape, box, seed)
    if boundaries != "zero" or periodicz:
        # a state that is not small at the walls, so the ghost values matter
        X, Y, Z = initial.mesh(shape, box)
        psi0 = psi0 + 3.0 * np.exp(1j * (0.7 * X - 0.4 * Y + 0.9 * Z + offset[0])) * \
            np.cos(0.3 * Y + offset[1])
    psi0 = psi0.astype(np.complex128 if prec == 64 else np.complex64)
    r = refbuild.run(d, psi0, prec=prec, nprocs=nprocs)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), psi0=psi0, psi=r["psi"],
                        t=r["t"], steps=steps, prec=prec, box=np.array(box),
                        harm_w=np.array(harm), omega_z=omega_z, gamma=gamma,
                        dt=np.complex128(dt), nprocs=nprocs, boundaries=boundaries,
                        periodicz=periodicz)
    return r


def case_euler(name="euler_f64_16", steps=20):
    """Reference imaginary time (src/run.f90:19-33) from its own 'soliton' initial
    state (setup_thomas_fermi, src/initial.f90:65-77,637-652): no restart file."""
    shape, box = (16, 16, 16), (16.0, 16.0, 16.0)
    harm = (4.0, 4.0, 4.0)   # TF_dense = 1 - pot stays positive over part of the box
    d = refbuild.build(name, prec=64, Nx=16, Ny=16, Nz=16, Lx=16.0, Ly=16.0, Lz=16.0,
                       nsteps=0, shots=steps, imaginary_time=True, imag_nsteps=steps,
                       imaginary_output=True, init_cond="soliton", dt_re=1e-3,
                       harm_wx=harm[0], harm_wy=harm[1], harm_wz=harm[2])
    r = refbuild.run(d, None, prec=64, nprocs=1, shape=shape, snapshot="imag_var0001.dat")
    np.savez_compressed(os.path.join(OUT, name + ".npz"), psi=r["psi"], steps=steps,
                        prec=64, box=np.array(box), harm_w=np.array(harm), dt=1e-3)


def case_c1(nprocs=4, steps=1000, seed=20261203):
    """Config 1 at full size: 64^3 FP64, 1000 steps, 4 ranks.  psi0 is regenerated from
    trapbec_amd.initial.config_c1(seed); stored: its SHA-256, every 4th point of the
    final psi, the z = 32 plane and the norms."""
    pr, psi0 = initial.config_c1(seed)
    kp = pr.kernel_params()
    d = refbuild.build("c1_f64", prec=64, Nx=64, Ny=64, Nz=64, Lx=16.0, Ly=16.0, Lz=16.0,
                       nsteps=steps, shots=steps, harm_wx=kp["harm_w"][0],
                       harm_wy=kp["harm_w"][1], harm_wz=kp["harm_w"][2])
    t0 = time.time()
    r = refbuild.run(d, psi0, prec=64, nprocs=nprocs)
    wall = time.time() - t0
    psi = r["psi"]
    np.savez_compressed(
        os.path.join(OUT, "c1_f64_64.npz"),
        psi0_sha256=hashlib.sha256(np.ascontiguousarray(psi0).tobytes()).hexdigest(),
        sub=psi[::4, ::4, ::4], plane=psi[:, :, 32], steps=steps, seed=seed,
        norm2=np.sum(np.abs(psi) ** 2), maxabs=np.abs(psi).max(), t=r["t"],
        nprocs=nprocs, wall_s=wall)
    return wall


def case_diag(name, prec, nprocs=1, shape=(16, 16, 16), box=(16.0, 16.0, 16.0), seed=13):
    """The output-time diagnostics (src/run.f90:64-66, src/diagnostic.f90:27-84): one
    real-time step from a seeded state with two vortex lines, then the reference's
    velocity (6th-order gradient of the unwrapped-nowhere phase) and quantum-pressure
    (2nd-order gradient of the 4th-order Laplacian of |psi|) records."""
    harm = tuple(L * L for L in box)
    d = refbuild.build(name, prec=prec, Nx=shape[0], Ny=shape[1], Nz=shape[2],
                       Lx=box[0], Ly=box[1], Lz=box[2], nsteps=1, shots=1,
                       harm_wx=harm[0], harm_wy=harm[1], harm_wz=harm[2],
                       vel_print=True, qpress_print=True)
    psi0 = small_ic(shape, box, seed)
    psi0 = psi0 * initial.vortex_line_factor(shape, box, -2.1, 1.7, -1, 1 / math.sqrt(10))
    psi0 = psi0.astype(np.complex128 if prec == 64 else np.complex64)
    r = refbuild.run(d, psi0, prec=prec, nprocs=nprocs)
    vel = refbuild.read_field(d, "vel", shape, nprocs, prec)
    qp = refbuild.read_field(d, "qpress", shape, nprocs, prec)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), psi=r["psi"], vel=vel, qpress=qp,
                        prec=prec, box=np.array(box), nprocs=nprocs)


def case_pot(name, prec, potential, shape=(24, 24, 24), box=(16.0, 16.0, 16.0), steps=20,
             t0=0.0, omega_z=0.0, potential_off=None, t_spindown=None, advective_vx=0.0,
             lbeam=False, wbeam=4.0, harm=None, seed=17, constant_ic=False, two_dim=False):
    """The potential select of get_rhs (src/finite_diff.f90:30-124) with the laser term,
    the switch-off times and the advective term: real-time RK3 from a seeded state
    restarted at t0 (so the time-dependent openings and tilts are exercised), or, for
    the mobius trap (whose Trap table the reference builds only on a fresh start,
    src/initial.f90:54-57,159-196), from the reference's own 'constant' state."""
    harm = tuple(L * L for L in box) if harm is None else harm
    d = refbuild.build(name, prec=prec, Nx=shape[0], Ny=shape[1], Nz=shape[2],
                       Lx=box[0], Ly=box[1], Lz=box[2], nsteps=steps, shots=steps,
                       omega_z=omega_z, potential=potential, potential_off=potential_off,
                       t_spindown=t_spindown, advective_vx=advective_vx, lbeam=lbeam,
                       wbeam=wbeam, harm_wx=harm[0], harm_wy=harm[1], harm_wz=harm[2],
                       two_dim=two_dim)
    cdt = np.complex128 if prec == 64 else np.complex64
    rdt = np.float64 if prec == 64 else np.float32
    if constant_ic:
        c = rdt(1) / np.sqrt(rdt(2))       # cmplx(1./sqrt(2.),1./sqrt(2.)), initial.f90:63
        psi0 = np.full(shape, c + 1j * c, dtype=cdt)
        r = refbuild.run(d, None, prec=prec, nprocs=1, shape=shape)
    else:
        psi0 = small_ic(shape, box, seed).astype(cdt)
        r = refbuild.run(d, psi0, prec=prec, nprocs=1, t0=t0)
    inf = np.inf
    np.savez_compressed(os.path.join(OUT, name + ".npz"), psi0=psi0, psi=r["psi"],
                        t=r["t"], t0=t0, steps=steps, prec=prec, box=np.array(box),
                        harm_w=np.array(harm), omega_z=omega_z, gamma=0.0,
                        dt=np.complex128(1e-3), nprocs=1, boundaries="zero",
                        periodicz=False, potential=potential,
                        potential_off=inf if potential_off is None else potential_off,
                        t_spindown=inf if t_spindown is None else t_spindown,
                        advective_vx=advective_vx, lbeam=lbeam, wbeam=wbeam, two_dim=two_dim)
    return r


def parse_shot_lines(stdout: str) -> np.ndarray:
    """The per-shot diagnostics line of program run (src/run.f90:59-63): list-directed
    `itime/shots, t, maxPsi, minPsi, meanPsi, intPsi2`, one row per shot.  List-directed
    output wraps long records over lines, so the values after the restart message are
    read as one token stream, six per row."""
    tail = stdout.split("restart, t=", 1)[1].split("\n", 1)[1]
    vals = [float(v) for v in tail.split()]
    assert len(vals) % 6 == 0, tail
    return np.array(vals, dtype=np.float64).reshape(-1, 6)


def case_info(name, prec, nprocs=1, shape=(12, 10, 8), box=(12.0, 10.0, 8.0), steps=20,
              shots=5, seed=23, keep_var=False):
    """Psi_info / Psi_norm (src/diagnostic.f90:7-25) as the reference prints them every
    `shots` steps (src/run.f90:59-63): real-time RK3 from a seeded state on a ragged grid;
    stored: psi0, the printed rows and the state of the first shot (var0001.dat).  keep_var
    also stores the raw per-rank snapshot records (src/output.f90:71-78) and dims.log
    (src/output.f90:48-58) for the gather test."""
    harm = tuple(L * L for L in box)
    d = refbuild.build(name, prec=prec, Nx=shape[0], Ny=shape[1], Nz=shape[2],
                       Lx=box[0], Ly=box[1], Lz=box[2], nsteps=steps, shots=shots,
                       harm_wx=harm[0], harm_wy=harm[1], harm_wz=harm[2])
    psi0 = small_ic(shape, box, seed).astype(np.complex128 if prec == 64 else np.complex64)
    r = refbuild.run(d, psi0, prec=prec, nprocs=nprocs)
    rows = parse_shot_lines(r["stdout"])
    assert rows.shape == (steps // shots, 6), (rows.shape, r["stdout"][-2000:])
    extra = {}
    if keep_var:
        with open(os.path.join(d, "data", "dims.log"), "rb") as fh:
            extra["raw_dims_log"] = np.frombuffer(fh.read(), np.uint8)
        for rk in range(nprocs):
            for shot in range(1, steps // shots + 1):
                fn = os.path.join(d, "data", f"proc{rk:03d}", f"var{shot:04d}.dat")
                with open(fn, "rb") as fh:
                    extra[f"raw_proc{rk:03d}_var{shot:04d}"] = np.frombuffer(fh.read(),
                                                                              np.uint8)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), psi0=psi0, psi=r["psi"],
                        rows=rows, steps=steps, shots=shots, prec=prec, box=np.array(box),
                        harm_w=np.array(harm), nprocs=nprocs, dt=1e-3, **extra)
    return r


def case_ic(name, prec, init_cond, shape=(20, 16, 12), box=(16.0, 14.0, 12.0),
            harm=(4.0, 4.0, 4.0), potential="harmonic", winding_number=1, ring_rad=4.0,
            add_ring=False, add_dipole=False, add_pcurrent=False, phase_imprint_on=False):
    """One of the reference's own initial states (src/initial.f90:58-156) with the
    insertions of src/run.f90:26,37-40, read back from var0001.dat after one step with
    dt = 0 (which leaves psi unchanged: w = i*0.5*0*rhs = 0).  phase_imprint runs in the
    reference's imaginary stage, one Euler step with the same dt = 0."""
    d = refbuild.build(name, prec=prec, Nx=shape[0], Ny=shape[1], Nz=shape[2],
                       Lx=box[0], Ly=box[1], Lz=box[2], nsteps=1, shots=1, dt_re=0.0,
                       dt_im=0.0, init_cond=init_cond, potential=potential,
                       harm_wx=harm[0], harm_wy=harm[1], harm_wz=harm[2],
                       winding_number=winding_number, ring_rad=ring_rad, add_ring=add_ring,
                       add_dipole=add_dipole, add_pcurrent=add_pcurrent,
                       phase_imprint_on=phase_imprint_on, imaginary_time=phase_imprint_on,
                       imag_nsteps=1)
    r = refbuild.run(d, None, prec=prec, nprocs=1, shape=shape)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), psi=r["psi"], prec=prec,
                        box=np.array(box), harm_w=np.array(harm), init_cond=init_cond,
                        potential=potential, winding_number=winding_number,
                        ring_rad=ring_rad, add_ring=add_ring, add_dipole=add_dipole,
                        add_pcurrent=add_pcurrent, phase_imprint_on=phase_imprint_on)
    return r


IC_CASES = {
    "ic_constant_f64": dict(prec=64, init_cond="constant"),
    "ic_soliton_f64": dict(prec=64, init_cond="soliton"),
    "ic_crow_f64": dict(prec=64, init_cond="crow", box=(16.0, 16.0, 12.0)),
    "ic_crow_f32": dict(prec=32, init_cond="crow", box=(16.0, 16.0, 12.0)),
    "ic_single_vortex_f64": dict(prec=64, init_cond="single_vortex", winding_number=2),
    "ic_sausage_f64": dict(prec=64, init_cond="sausage"),
    "ic_leap_frog_f64": dict(prec=64, init_cond="leap_frog", box=(16.0, 32.0, 12.0)),
    "ic_line_of_lines_f64": dict(prec=64, init_cond="line_of_lines", ring_rad=3.0),
    "ic_linked_rings_f64": dict(prec=64, init_cond="linked_rings", ring_rad=3.0),
    "ic_lattice_f64": dict(prec=64, init_cond="lattice"),
    "ic_lattice_f32": dict(prec=32, init_cond="lattice"),
    "ic_lattice2_f64": dict(prec=64, init_cond="lattice2", box=(40.0, 40.0, 40.0),
                            harm=(1.0, 1.0, 1.0)),
    "ic_bundle_f64": dict(prec=64, init_cond="bundle", box=(80.0, 80.0, 80.0),
                          harm=(1.0, 1.0, 1.0)),
    "ins_ring_f64": dict(prec=64, init_cond="constant", add_ring=True, ring_rad=3.0),
    "ins_dipole_f64": dict(prec=64, init_cond="constant", add_dipole=True,
                           box=(48.0, 24.0, 12.0)),
    "ins_pcurrent_barrier_f64": dict(prec=64, init_cond="constant", add_pcurrent=True,
                                     potential="barrier"),
    "ins_pcurrent_mobius_f64": dict(prec=64, init_cond="constant", add_pcurrent=True,
                                    potential="mobius"),
    "ins_pcurrent_ring_f64": dict(prec=64, init_cond="constant", add_pcurrent=True,
                                  potential="harmonic_ring"),
    "ins_pcurrent_ring2_f64": dict(prec=64, init_cond="constant", add_pcurrent=True,
                                   potential="harmonic_ring2"),
    "ins_pcurrent_channel_f32": dict(prec=32, init_cond="constant", add_pcurrent=True,
                                     potential="channel"),
    "ins_phase_imprint_f64": dict(prec=64, init_cond="constant", phase_imprint_on=True,
                                  winding_number=2),
}


CASES = {
    "rk3_f64_16": lambda: case_small("rk3_f64_16", 64),
    "rk3_f64_16_rot": lambda: case_small("rk3_f64_16_rot", 64, omega_z=-1.4),
    "rk3_f32_16": lambda: case_small("rk3_f32_16", 32),
    "rk3_f32_16_rot": lambda: case_small("rk3_f32_16_rot", 32, omega_z=-1.4),
    "rk3_f64_16_gamma": lambda: case_small("rk3_f64_16_gamma", 64, gamma=0.05),
    "rk3_f64_16_imag": lambda: case_small("rk3_f64_16_imag", 64, dt=-1e-3j),
    "rk3_f64_aniso": lambda: case_small("rk3_f64_aniso", 64, shape=(16, 12, 20),
                                        box=(16.0, 12.0, 10.0), omega_z=-0.8),
    "rk3_f64_16_np4": lambda: case_small("rk3_f64_16_np4", 64, omega_z=-1.4, nprocs=4),
    "rk3_f64_16_solid": lambda: case_small("rk3_f64_16_solid", 64, omega_z=-1.4,
                                           boundaries="solid"),
    "rk3_f32_16_solid": lambda: case_small("rk3_f32_16_solid", 32, boundaries="solid"),
    "rk3_f64_16_periodic": lambda: case_small("rk3_f64_16_periodic", 64, omega_z=-1.4,
                                              boundaries="periodic"),
    "rk3_f64_16_periodic_np4": lambda: case_small("rk3_f64_16_periodic_np4", 64,
                                                  omega_z=-1.4, boundaries="periodic",
                                                  nprocs=4),
    "rk3_f64_aniso_periodicz": lambda: case_small("rk3_f64_aniso_periodicz", 64,
                                                  shape=(16, 12, 20), box=(16.0, 12.0, 10.0),
                                                  omega_z=-0.8, periodicz=True),
    "rk3_f64_16_solid_np4": lambda: case_small("rk3_f64_16_solid_np4", 64, omega_z=-1.4,
                                               boundaries="solid", nprocs=4),
    "info_f64_ragged": lambda: case_info("info_f64_ragged", 64),
    "info_f32_ragged": lambda: case_info("info_f32_ragged", 32),
    "info_f64_ragged_np4": lambda: case_info("info_f64_ragged_np4", 64, nprocs=4,
                                             keep_var=True),
    "info_f32_ragged_np4": lambda: case_info("info_f32_ragged_np4", 32, nprocs=4),
    "euler_f64_16": case_euler,
    "diag_f64_16": lambda: case_diag("diag_f64_16", 64),
    "diag_f32_16": lambda: case_diag("diag_f32_16", 32),
    "diag_f64_16_np4": lambda: case_diag("diag_f64_16_np4", 64, nprocs=4),
    "c1_f64_64": case_c1,
    "pot_ring_f64": lambda: case_pot("pot_ring_f64", 64, "harmonic_ring", t0=49.99),
    "pot_ring2_f32": lambda: case_pot("pot_ring2_f32", 32, "harmonic_ring2", t0=49.995),
    "pot_barrier_f64": lambda: case_pot("pot_barrier_f64", 64, "barrier",
                                        potential_off=0.0105),
    "pot_laser_f64": lambda: case_pot("pot_laser_f64", 64, "zero", shape=(48, 16, 16),
                                      box=(160.0, 32.0, 16.0), lbeam=True, wbeam=8.0),
    "pot_box_f64": lambda: case_pot("pot_box_f64", 64, "box_trap", t0=15.0, omega_z=-1.4,
                                    t_spindown=15.0105),
    "pot_channel_f32": lambda: case_pot("pot_channel_f32", 32, "channel", shape=(20, 20, 20),
                                        advective_vx=0.3),
    "pot_quartic_f32": lambda: case_pot("pot_quartic_f32", 32, "quartic", shape=(16, 16, 16)),
    "pot_quartic_f64": lambda: case_pot("pot_quartic_f64", 64, "quartic", shape=(16, 16, 16)),
    "pot_sphere_f64": lambda: case_pot("pot_sphere_f64", 64, "sphere"),
    "pot_sphere_f32": lambda: case_pot("pot_sphere_f32", 32, "sphere"),
    "pot_mobius_f64": lambda: case_pot("pot_mobius_f64", 64, "mobius", shape=(32, 32, 16),
                                       box=(64.0, 64.0, 16.0), steps=10, constant_ic=True),
    "pot_harm_switch_f64": lambda: case_pot("pot_harm_switch_f64", 64, "harmonic",
                                            shape=(16, 16, 16), omega_z=-1.4,
                                            potential_off=0.0105, t_spindown=0.0155),
    "pot_2d_f64": lambda: case_pot("pot_2d_f64", 64, "harmonic", shape=(32, 24, 1),
                                   box=(16.0, 12.0, 1.0), harm=(256.0, 144.0, 0.0),
                                   omega_z=-1.4, two_dim=True),
    "pot_2d_f32": lambda: case_pot("pot_2d_f32", 32, "barrier", shape=(32, 24, 1),
                                   box=(16.0, 12.0, 1.0), two_dim=True,
                                   potential_off=0.0155),
    "pot_advective_f64": lambda: case_pot("pot_advective_f64", 64, "harmonic",
                                          shape=(16, 16, 16), omega_z=-1.4,
                                          advective_vx=0.5),
}

for _n, _kw in IC_CASES.items():
    CASES[_n] = (lambda n, kw: (lambda: case_ic(n, **kw)))(_n, _kw)

if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for n in names:
        t0 = time.time()
        CASES[n]()
        print(f"{n}: {time.time() - t0:.1f}s", flush=True)
