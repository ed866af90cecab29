"""The CPU oracle, pinned: against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from the compiled reference), and against closed forms
chosen so that a dropped term, a wrong sign or a wrong coefficient fails."""
import glob
import hashlib
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, oracle
import refic_draws
from trapbec_amd import initial

SMALL = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "rk3_*.npz")))


def _stepper_from_fixture(f):
    prec = int(f["prec"])
    box = tuple(float(b) for b in f["box"])
    bc = str(f["boundaries"]) if "boundaries" in f else "zero"
    pz = bool(f["periodicz"]) if "periodicz" in f else False
    def opt(key, default, conv=float):
        if key not in f:
            return default
        v = conv(f[key])
        return None if (conv is float and np.isinf(v)) else v
    p = oracle.params(f["psi0"].shape, box, tuple(float(h) for h in f["harm_w"]),
                      omega_z=float(f["omega_z"]), dt=complex(f["dt"]),
                      gamma=float(f["gamma"]), boundaries=bc, periodicz=pz,
                      potential=opt("potential", "harmonic", str),
                      potential_off=opt("potential_off", None),
                      t_spindown=opt("t_spindown", None),
                      advective_vx=opt("advective_vx", 0.0),
                      lbeam=opt("lbeam", False, bool), wbeam=opt("wbeam", 4.0),
                      two_dim=opt("two_dim", False, bool))
    s = oracle.Stepper(p, f["psi0"], prec=prec, t=opt("t0", 0.0))
    if "potential" in f and str(f["potential"]) == "mobius":
        s.set_trap(s.mobius_trap())
    return s, prec


@pytest.mark.parametrize("name", SMALL)
def test_matches_reference_bitwise(name):
    """RK3 (src/run.f90:42-55) over 20 steps equals the reference build bit for bit:
    FP64 and FP32, rotation on/off, dissipation, complex (imaginary) dt, an anisotropic
    grid, and a 4-rank run (decomposition invariance)."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    s, prec = _stepper_from_fixture(f)
    s.rk3(int(f["steps"]))
    got, ref = s.interior, f["psi"]
    assert got.dtype == ref.dtype
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), \
        f"{np.sum(got != ref)} points differ, max {np.abs(got - ref).max()}"
    assert s.t == pytest.approx(float(f["t"]), rel=1e-6 if prec == 32 else 1e-15)


POT = ["pot_ring_f64", "pot_ring2_f32", "pot_barrier_f64", "pot_laser_f64", "pot_box_f64",
       "pot_channel_f32", "pot_quartic_f32", "pot_quartic_f64", "pot_sphere_f64",
       "pot_sphere_f32", "pot_mobius_f64", "pot_harm_switch_f64", "pot_advective_f64",
       "pot_2d_f64", "pot_2d_f32"]


@pytest.mark.parametrize("name", POT)
def test_potentials_match_reference_bitwise(name):
    """The potential select (src/finite_diff.f90:30-124): ring openings closing at
    t = 50, the barrier switched off mid-run, the laser beam, the box trap's sin(0.1 t)
    tilt with the frame term spun down mid-run, the channel with the advective term,
    quartic (r**8), sphere, the mobius Trap table (src/initial.f90:160-196) and the
    harmonic trap with both switch-off times, two_dim on a 32 x 24 x 1 grid: 20 (mobius
    10) RK3 steps equal the reference build bit for bit."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    s, prec = _stepper_from_fixture(f)
    s.rk3(int(f["steps"]))
    got, ref = s.interior, f["psi"]
    assert got.dtype == ref.dtype
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), \
        f"{np.sum(got != ref)} points differ, max {np.abs(got - ref).max()}"
    assert s.t == pytest.approx(float(f["t"]), rel=1e-6 if prec == 32 else 1e-15)


def test_decomposition_invariance():
    """The 4-rank reference run equals the 1-rank run bit for bit (same psi0, rotation)."""
    a = np.load(os.path.join(GOLDEN, "rk3_f64_16_rot.npz"))
    b = np.load(os.path.join(GOLDEN, "rk3_f64_16_np4.npz"))
    assert np.array_equal(a["psi0"], b["psi0"])
    assert np.array_equal(a["psi"], b["psi"])


def test_reference_euler_imaginary_time():
    """Reference imaginary time (src/run.f90:19-33) from its own 'soliton' state
    (setup_thomas_fermi, src/initial.f90:637-652 + the sign flip at :65-77)."""
    f = np.load(os.path.join(GOLDEN, "euler_f64_16.npz"))
    box = tuple(float(b) for b in f["box"])
    harm = tuple(float(h) for h in f["harm_w"])
    psi0 = initial.reference_soliton((16, 16, 16), box, harm)
    p = oracle.params((16, 16, 16), box, harm, dt=float(f["dt"]))
    s = oracle.Stepper(p, psi0)
    s.euler(int(f["steps"]))
    ref = f["psi"]
    # exp(i*pi) of the reference's 12-digit pi comes from the Fortran runtime, so the
    # initial state agrees to rounding, not bit for bit.
    assert np.abs(s.interior - ref).max() <= 1e-14 * np.abs(ref).max()


def test_c1_full_size():
    """Config 1 at full size: 64^3, FP64, 1000 steps against the 4-rank reference run."""
    path = os.path.join(GOLDEN, "c1_f64_64.npz")
    if not os.path.exists(path):
        pytest.skip("c1 fixture not generated")
    f = np.load(path)
    pr, psi0 = initial.config_c1(int(f["seed"]))
    assert hashlib.sha256(np.ascontiguousarray(psi0).tobytes()).hexdigest() == \
        str(f["psi0_sha256"])
    kp = pr.kernel_params()
    p = oracle.params(pr.shape, pr.box, kp["harm_w"], dt=1e-3)
    s = oracle.Stepper(p, psi0)
    s.rk3(int(f["steps"]))
    psi = s.interior
    assert np.array_equal(psi[::4, ::4, ::4], f["sub"])
    assert np.array_equal(psi[:, :, 32], f["plane"])


# ----------------------------------------------------------------- closed forms ----

def _lap_only(shape, box):
    """Parameters that reduce get_rhs to the Laplacian (no trap, no constant, no g)."""
    return oracle.params(shape, box, (0.0, 0.0, 0.0), c0=0.0, gnl=0.0)


def test_second_derivative_exact_on_polynomials():
    """The 6th-order D2 is exact on polynomials of degree <= 7; points >= 3 cells from
    the zero ghosts see only the polynomial."""
    n, L = 20, 5.0
    X, Y, Z = initial.mesh((n, n, n), (L, L, L))
    f = (0.3 * X ** 7 - X ** 3 * Y ** 2 + 2 * Z ** 5 + 1j * (Y ** 6 - X * Z))
    lap = (0.3 * 42 * X ** 5 - 6 * X * Y ** 2 + 40 * Z ** 3 + 1j * 30 * Y ** 4) \
        + (-2 * X ** 3)
    s = oracle.Stepper(_lap_only((n, n, n), (L, L, L)), f)
    r = s.get_rhs()
    sl = slice(3, n - 3)
    err = np.abs(r[sl, sl, sl] - lap[sl, sl, sl]).max()
    assert err <= 1e-9 * np.abs(lap).max()


def test_second_derivative_modified_wavenumber():
    """Plane wave e^{ikx}: D2 = (4cos3t - 54cos2t + 540cos t - 490)/(180h^2), t = kh."""
    n, L = 24, 6.0
    h = L / n
    k = 2.3
    X, _, _ = initial.mesh((n, n, n), (L, L, L))
    f = np.exp(1j * k * X)
    t = k * h
    sym = (4 * math.cos(3 * t) - 54 * math.cos(2 * t) + 540 * math.cos(t) - 490) / (180 * h * h)
    s = oracle.Stepper(_lap_only((n, n, n), (L, L, L)), f)
    r = s.get_rhs()
    # interior in x; y and z see a constant, whose D2 is zero away from the ghosts
    sl = slice(3, n - 3)
    err = np.abs(r[sl, sl, sl] - sym * f[sl, sl, sl]).max()
    assert err < 1e-10


def test_frame_term_first_derivative():
    """om*i*(x d_y - y d_x) psi: exact for polynomials of degree <= 6 (D1 is 6th order);
    for psi = (x + i y)^3, (x d_y - y d_x) psi = 3 i psi."""
    n, L = 20, 4.0
    X, Y, Z = initial.mesh((n, n, n), (L, L, L))
    f = (X + 1j * Y) ** 3 * (1 + 0.1 * Z)
    om = 0.7
    base = oracle.Stepper(_lap_only((n, n, n), (L, L, L)), f).get_rhs()
    p = _lap_only((n, n, n), (L, L, L))
    p.omega_z = om
    rot = oracle.Stepper(p, f).get_rhs()
    sl = slice(3, n - 3)
    frame = (rot - base)[sl, sl, sl]
    expect = om * 1j * (3j * f)[sl, sl, sl]
    assert np.abs(frame - expect).max() <= 1e-11 * np.abs(expect).max()


def test_potential_and_nonlinearity_single_point():
    """A single nonzero point: rhs = 3*(-490/(180 h^2)) psi + (c0 - gnl|psi|^2 - pot) psi
    with pot = sum harm_w (x/L)^2 at that cell centre; its 6 axis neighbours at distance
    1..3 see the D2 weights 270, -27, 2 over 180 h^2."""
    n, L = 12, 6.0
    h = L / n
    f = np.zeros((n, n, n), complex)
    i, j, k = 4, 7, 5
    f[i, j, k] = 0.8 - 0.3j
    harm = (3.0, 5.0, 7.0)
    p = oracle.params((n, n, n), (L, L, L), harm, c0=0.9, gnl=1.7)
    r = oracle.Stepper(p, f).get_rhs()
    x = initial.coords(n, L)
    pot = harm[0] * (x[i] / L) ** 2 + harm[1] * (x[j] / L) ** 2 + harm[2] * (x[k] / L) ** 2
    psi = f[i, j, k]
    want = 3 * (-490 / (180 * h * h)) * psi + (0.9 - 1.7 * abs(psi) ** 2 - pot) * psi
    assert abs(r[i, j, k] - want) < 1e-12 * abs(want)
    for d, wgt in ((1, 270), (2, -27), (3, 2)):
        assert abs(r[i + d, j, k] - wgt / (180 * h * h) * psi) < 1e-12 * abs(psi) * 1e3
        assert abs(r[i, j, k - d] - wgt / (180 * h * h) * psi) < 1e-12 * abs(psi) * 1e3
    assert r[i + 4, j, k] == 0 and r[i + 1, j + 1, k] == 0


def test_rk3_linear_stability_polynomial():
    """For a linear right-hand side (gnl = 0) one RK3 step is R(A) psi with
    R(z) = 1 + z + z^2/2 + z^3/6 and A psi = (i dt/2) rhs(psi) (src/run.f90:46-54),
    so the A = (0,-2/3,-1), B = (1/3,1,1/2) coefficients are pinned."""
    n, L = 12, 6.0
    rng = np.random.default_rng(5)
    f = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    dt = 2e-3 - 1e-3j
    p = oracle.params((n, n, n), (L, L, L), (2.0, 3.0, 4.0), omega_z=0.6, c0=0.5,
                      gnl=0.0, dt=dt)

    def A(v):
        return 0.5j * dt * oracle.Stepper(p, v).get_rhs()

    a1 = A(f)
    a2 = A(a1)
    a3 = A(a2)
    want = f + a1 + a2 / 2 + a3 / 6
    s = oracle.Stepper(p, f)
    s.rk3(1)
    assert np.abs(s.interior - want).max() < 1e-12 * np.abs(f).max()


def test_rk3_second_order_nonlinear():
    """The scheme's b.c^2 = 5/18 != 1/3 makes it 2nd order on the nonlinear GPE: halving
    dt shrinks the one-step-to-T error by ~4 (not ~8)."""
    n, L = 12, 6.0
    f = initial.tf_profile((n, n, n), (L, L, L), (1, 1, 1), 1.0, mu=3.0)[0] + 0.2
    f = f * (1 + 0.3j)

    def run(dt, steps):
        p = oracle.params((n, n, n), (L, L, L), (L * L,) * 3, c0=0.0, gnl=40.0, dt=dt)
        s = oracle.Stepper(p, f)
        s.rk3(steps)
        return s.interior

    ref = run(0.00025, 64)
    e1 = np.abs(run(0.004, 4) - ref).max()
    e2 = np.abs(run(0.002, 8) - ref).max()
    rate = math.log2(e1 / e2)
    assert 1.6 < rate < 2.6, rate


# ------------------------------------ FP64 quantities with no reference code -------

def _gaussian(shape, box, omega):
    X, Y, Z = initial.mesh(shape, box)
    wx, wy, wz = omega
    return np.exp(-0.5 * (wx * X * X + wy * Y * Y + wz * Z * Z)).astype(complex)


def _native_params(shape, box, omega, g, Omega, dt=1e-3):
    pr = initial.Problem(shape, box, omega, g, Omega, dt)
    kp = pr.kernel_params()
    return oracle.params(shape, box, kp["harm_w"], omega_z=kp["omega_r"], c0=kp["c0"],
                         gnl=kp["gnl"], dt=kp["dt"])


def test_energy_harmonic_ground_state():
    """g = 0: the Gaussian is the ground state, E = mu = (wx + wy + wz)/2, residual ~ 0."""
    shape, box, om = (40, 40, 40), (12.0, 12.0, 12.0), (1.0, 1.3, 0.7)
    s = oracle.Stepper(_native_params(shape, box, om, 0.0, 0.0), _gaussian(shape, box, om))
    e = s.energy()
    # 6th-order discretisation error at h = 0.3 is ~1e-5
    assert e["energy"] / e["norm"] == pytest.approx(0.5 * sum(om), rel=5e-5)
    assert e["mu"] == pytest.approx(0.5 * sum(om), rel=5e-5)
    assert e["residual"] < 1e-3   # an exact eigenstate up to the same discretisation


def test_energy_rotating_frame_sign():
    """psi = (x + i y) exp(-r^2/2) in an isotropic trap has L_z = +1 and E = 5/2 without
    rotation, so E = 5/2 - Omega in the frame rotating at Omega (H - Omega L_z)."""
    shape, box = (40, 40, 40), (12.0, 12.0, 12.0)
    X, Y, _ = initial.mesh(shape, box)
    f = (X + 1j * Y) * _gaussian(shape, box, (1, 1, 1))
    for Om in (0.0, 0.4):
        s = oracle.Stepper(_native_params(shape, box, (1, 1, 1), 0.0, Om), f)
        e = s.energy()
        assert e["energy"] / e["norm"] == pytest.approx(2.5 - Om, rel=1e-4)
        assert e["residual"] < 1e-3


def test_energy_interaction_scaling():
    """E(lam psi) = lam^2 (E_kin + E_pot) + lam^4 (g/2) sum |psi|^4 dV."""
    shape, box, g = (24, 24, 24), (10.0, 10.0, 10.0), 37.0
    f = _gaussian(shape, box, (1, 1, 1)) * np.exp(0.3j * initial.mesh(shape, box)[0])
    p = _native_params(shape, box, (1, 1, 1), g, 0.2)
    e1 = oracle.Stepper(p, f).energy()["energy"]
    e2 = oracle.Stepper(p, 2 * f).energy()["energy"]
    dV = (10.0 / 24) ** 3
    eint = 0.5 * g * np.sum(np.abs(f) ** 4) * dV
    assert (e2 - 4 * e1) / 12 == pytest.approx(eint, rel=1e-10)


def test_imaginary_time_virial():
    """Renormalised imaginary-time RK3 converges to the interacting ground state, whose
    energy satisfies mu*N = E + E_int and the virial 2E_kin - 2E_pot + 3E_int = 0."""
    shape, box, om, g = (24, 24, 24), (12.0, 12.0, 12.0), (1.0, 1.0, 1.0), 20.0
    p = _native_params(shape, box, om, g, 0.0, dt=-1e-2j)
    s = oracle.Stepper(p, _gaussian(shape, box, (0.8, 0.8, 0.8)))
    s.imag_rk3(2500, 1.0)
    e = s.energy()
    assert e["norm"] == pytest.approx(1.0, rel=1e-12)
    assert e["residual"] < 1e-6
    psi = s.interior
    dV = (12.0 / 24) ** 3
    eint = 0.5 * g * np.sum(np.abs(psi) ** 4) * dV
    V = initial.trap_potential(shape, box, om)
    epot = np.sum(V * np.abs(psi) ** 2) * dV
    ekin = e["energy"] - epot - eint
    assert e["mu"] == pytest.approx(e["energy"] + eint, rel=1e-8)
    assert abs(2 * ekin - 2 * epot + 3 * eint) < 2e-3 * e["energy"]


def test_vortex_plaquettes_lines():
    """Imprinted straight lines give one plaquette per z-plane with the imprinted charge
    at the cell containing the line."""
    shape, box = (32, 32, 8), (16.0, 16.0, 4.0)
    lines = [(-3.1, 2.2, 1), (2.6, -1.3, -1), (0.4, 4.3, 1)]
    f = initial.tf_profile(shape, box, (0.3, 0.3, 0.3), 1.0, mu=10.0)[0]
    for x0, y0, q in lines:
        f = f * initial.vortex_line_factor(shape, box, x0, y0, q, 0.5)
    s = oracle.Stepper(_lap_only(shape, box), f)
    v = s.vortices(1e-6)
    assert len(v) == len(lines) * shape[2]
    h = 16.0 / 32
    x = initial.coords(32, 16.0)
    for x0, y0, q in lines:
        i = int(np.searchsorted(x, x0)) - 1
        j = int(np.searchsorted(x, y0)) - 1
        hits = v[(v[:, 0] == i) & (v[:, 1] == j)]
        assert len(hits) == shape[2] and np.all(hits[:, 3] == q), (x0, y0, h)


def test_vortex_plaquettes_ring_orientations():
    """A vortex ring in the z=0 plane pierces x-normal and y-normal plaquettes (two
    opposite charges on each axis through the centre) and no z-normal one."""
    shape, box = (40, 40, 40), (12.0, 12.0, 12.0)
    f = initial.tf_profile(shape, box, (0.3, 0.3, 0.3), 1.0, mu=10.0)[0] * \
        initial.vortex_ring_factor(shape, box, 3.0, 0.0, 0.4)
    s = oracle.Stepper(_lap_only(shape, box), f)
    assert len(s.vortices(1e-6, 0)) == 0
    vx = s.vortices(1e-6, 1)       # x-normal plaquettes: the ring cuts the x = const planes
    vy = s.vortices(1e-6, 2)
    assert len(vx) > 0 and len(vy) > 0
    assert set(np.unique(vx[:, 3])) == {-1, 1} and set(np.unique(vy[:, 3])) == {-1, 1}
    x = initial.coords(40, 12.0)
    # every x-normal hit sits on the ring: rho(y, z=0) ~ 3 within a cell
    yy = x[vx[:, 1]] + 0.15
    xx = x[vx[:, 0]]
    rho = np.sqrt(xx ** 2 + yy ** 2)
    assert np.all(np.abs(rho - 3.0) < 0.5)


@pytest.mark.parametrize("name", ["diag_f64_16", "diag_f32_16"])
def test_diagnostics_match_reference_bitwise(name):
    """get_phase's velocity and get_qpressure's field (src/diagnostic.f90:27-83), as the
    reference wrote them (vel01.dat, qpress01.dat), bit for bit from the same psi."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    prec = int(f["prec"])
    box = tuple(float(b) for b in f["box"])
    s = oracle.Stepper(oracle.params(f["psi"].shape, box, tuple(L * L for L in box)),
                       f["psi"], prec=prec)
    assert np.array_equal(s.velocity(), f["vel"])
    assert np.array_equal(s.qpressure(), f["qpress"])


def test_diagnostics_four_ranks():
    """On 2 x 2 ranks the reference's velocity equals its one-rank field.  Its quantum
    pressure does not: ghost_comm fills face ghosts only (src/ghost.f90:27-192), and
    the Laplacian on the one-cell-extended box reads the unfilled corner ghosts next
    to inner rank boundaries.  Rebuilding each rank's padded block that way (faces from
    the neighbours, corners zero) reproduces the 4-rank file bit for bit; the single
    GPU domain has no such corners and matches the one-rank file."""
    one = np.load(os.path.join(GOLDEN, "diag_f64_16.npz"))
    four = np.load(os.path.join(GOLDEN, "diag_f64_16_np4.npz"))
    assert np.array_equal(four["psi"], one["psi"])
    assert np.array_equal(four["vel"], one["vel"])
    assert not np.array_equal(four["qpress"], one["qpress"])
    psi = one["psi"]
    nx, ny, nz = psi.shape
    box = tuple(float(b) for b in one["box"])
    full = oracle.pad(psi)
    mx, my = nx // 2, ny // 2
    qp = np.zeros_like(one["qpress"])
    for cx in range(2):
        for cy in range(2):
            x0, y0 = cx * mx, cy * my
            blk = np.zeros((mx + 6, my + 6, nz + 6), dtype=np.complex128)
            # interior plus x faces (all interior y) plus y faces (all interior x)
            blk[:, 3:-3, :] = full[x0:x0 + mx + 6, y0 + 3:y0 + 3 + my, :]
            blk[3:-3, :, :] = full[x0 + 3:x0 + 3 + mx, y0:y0 + my + 6, :]
            p = oracle.params((nx, ny, nz), box, tuple(L * L for L in box), dims=(2, 2),
                              coords=(cx, cy))
            s = oracle.Stepper(p, np.zeros((mx, my, nz)))
            s.psi = blk
            qp[:, x0:x0 + mx, y0:y0 + my, :] = s.qpressure()
    assert np.array_equal(qp, four["qpress"])


def test_diagnostics_closed_forms():
    """Velocity of a plane wave exp(i k x) is k away from the phase jumps (6th order);
    the quantum-pressure gradient of a Gaussian |psi| matches its analytic form to the
    stencils' order."""
    n, L = 48, 12.0
    shape = (n, n, n)
    X, Y, Z = initial.mesh(shape, (L, L, L))
    kx = 2 * math.pi / L * 2
    psi = np.exp(1j * kx * X) * np.exp(-0.02 * (X ** 2 + Y ** 2 + Z ** 2))
    s = oracle.Stepper(oracle.params(shape, L, (1.0, 1.0, 1.0)), psi)
    v = s.velocity()
    ph = np.angle(psi)
    h = L / n
    smooth = np.ones(shape, bool)
    for sh in range(-3, 4):   # no 2*pi jump inside the x stencil
        smooth &= np.abs(np.roll(ph, sh, 0) - ph) < math.pi
    smooth[:3] = smooth[-3:] = False
    assert np.abs(v[0][smooth] - kx).max() < 1e-6
    assert np.abs(v[1][3:-3, 3:-3, 3:-3]).max() < 1e-10   # phase independent of y
    a = 0.3
    errs = []
    for m in (48, 96):
        Xm, Ym, Zm = initial.mesh((m, m, m), (L, L, L))
        r2 = Xm ** 2 + Ym ** 2 + Zm ** 2
        rho = np.exp(-a * r2)                                  # |psi| itself
        s = oracle.Stepper(oracle.params((m, m, m), L, (1.0, 1.0, 1.0)), rho.astype(complex))
        q = s.qpressure()
        # lap(rho) = (4 a^2 r^2 - 6 a) rho; its x derivative:
        dlap = (8 * a * a * Xm - 2 * a * Xm * (4 * a * a * r2 - 6 * a)) * rho
        inner = (np.abs(Xm) < 3) & (np.abs(Ym) < 3) & (np.abs(Zm) < 3)
        errs.append(np.abs(q[0][inner] - dlap[inner]).max() / np.abs(dlap[inner]).max())
    assert errs[1] < 1e-2
    assert 3.5 < errs[0] / errs[1] < 4.5          # second order (low_first_deriv)
    assert h < 0.3


def test_imprint_vortex_closed_forms():
    """force_random_points' imprint (src/force.f90:32-53) on a uniform state: the
    magnitude is Berloff's Pade approximant of the GP vortex core, f^2 = r^2 (a1 + a2 r^2)
    / (1 + b1 r^2 + a2 r^4) with the published a1 = 0.34375, b1 = 1/3, a2 = 11/384
    (N. G. Berloff, J. Phys. A 37, 1617 (2004)); the phase winds once (sense dir) around
    exactly one plaquette per z-plane at the centre; +1 then -1 at the same centre leaves
    the real factor f^2 and no phase."""
    shape, box = (24, 20, 8), (16.0, 14.0, 4.0)
    x0, y0 = 1.3, -2.1
    psi0 = np.ones(shape, dtype=np.complex128)
    p = oracle.params(shape, box, (1.0, 1.0, 1.0))
    s = oracle.Stepper(p, psi0)
    s.imprint_vortex(x0, y0, 1.0)
    X, Y, _ = initial.mesh(shape, box)
    r2 = (X - x0) ** 2 + (Y - y0) ** 2
    a1, b1, a2 = 0.34375, 1.0 / 3.0, 11.0 / 384.0
    f = np.sqrt(r2 * (a1 + a2 * r2) / (1 + b1 * r2 + a2 * r2 * r2))
    got = s.interior
    assert np.abs(np.abs(got) - f).max() < 1e-13
    v = s.vortices(1e-6, 0)
    assert len(v) == shape[2] and set(v[:, 3]) == {1}
    h = (box[0] / shape[0], box[1] / shape[1])
    cx = (v[:, 0] + 1) * h[0] - box[0] / 2          # plaquette centre (corners at cells)
    cy = (v[:, 1] + 1) * h[1] - box[1] / 2
    assert np.all(np.abs(cx - x0) < h[0]) and np.all(np.abs(cy - y0) < h[1])
    s.imprint_vortex(x0, y0, -1.0)
    back = s.interior
    assert np.abs(back.imag).max() < 1e-14
    assert np.abs(back.real - f * f).max() < 1e-13
    s2 = oracle.Stepper(p, psi0)
    s2.imprint_vortex(x0, y0, -1.0)
    assert set(s2.vortices(1e-6, 0)[:, 3]) == {-1}


IC_FIXTURES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "ic_*.npz")))
INS_FIXTURES = sorted(os.path.basename(p)[:-4]
                      for p in glob.glob(os.path.join(GOLDEN, "ins_*.npz")))


def oracle_ic(f):
    """The oracle's restatement of a fixture's initial state and insertions."""
    prec = int(f["prec"])
    box = tuple(float(b) for b in f["box"])
    shape = f["psi"].shape
    kind = str(f["init_cond"])
    p = oracle.params(shape, box, tuple(float(h) for h in f["harm_w"]),
                      potential=str(f["potential"]))
    s = oracle.Stepper(p, np.zeros(shape, f["psi"].dtype), prec=prec)
    wn, rr = int(f["winding_number"]), float(f["ring_rad"])
    prm = None
    if kind == "bundle":
        prm = refic_draws.bundle_centres(prec)
    elif kind == "soliton":
        prm = refic_draws.exp_i_pi(prec)
    s.reference_state(kind, wn, rr, prm)
    if bool(f["phase_imprint_on"]):
        s.reference_insert("phase_imprint", wn, rr)
    if bool(f["add_pcurrent"]):
        if str(f["potential"]) == "mobius":
            s.set_trap(s.mobius_trap())
        s.reference_insert("pcurrent", wn, rr)
    if bool(f["add_ring"]):
        s.reference_insert("ring", wn, rr)
    if bool(f["add_dipole"]):
        s.reference_insert("dipole", wn, rr)
    return s.interior


@pytest.mark.parametrize("name", IC_FIXTURES + INS_FIXTURES)
def test_initial_states_match_reference(name):
    """The reference's own init_cond states (constant, soliton, crow, single_vortex with
    winding 2, sausage, leap_frog, line_of_lines, linked_rings, lattice, lattice2,
    bundle) and insertions (add_ring, add_dipole, add_pcurrent for each potential that
    defines one, phase_imprint), from var0001.dat of a dt = 0 run: equal as values, except
    the soliton's exp(eye*pi), which amdflang folds at compile time (rounding).  Values,
    not bytes: amdflang multiplies real x complex as the full complex product
    (d, 0)*(c, s), whose d*s + 0*c turns a -0 into +0 where the oracle keeps d*s; such
    signed zeros (3 points of the bundle state, in its zero-density region) compare
    equal and enter no later arithmetic."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    got, ref = oracle_ic(f), f["psi"]
    assert got.dtype == ref.dtype
    if str(f["init_cond"]) == "soliton":
        assert np.abs(got - ref).max() <= 4 * np.finfo(ref.real.dtype).eps * np.abs(ref).max()
        return
    assert np.array_equal(got, ref), \
        f"{np.sum(got != ref)} points differ, max {np.abs(got - ref).max()}"


def test_random_phase_is_the_drawn_plane_waves():
    """setup_quench (src/initial.f90:346-375) with a flat Thomas-Fermi factor (harm_w = 0):
    the state is the sum of the 20 drawn plane waves exp(2 pi i k.x / L), so its discrete
    Fourier transform on the cell-centred 40^3 grid holds each drawn amplitude at its wave
    vector (duplicates add) and nothing else."""
    n, L = 40, 10.0
    prm = refic_draws.random_phase(7, 64)
    s = oracle.Stepper(oracle.params((n, n, n), L, (0.0, 0.0, 0.0)),
                       np.zeros((n, n, n), np.complex128))
    s.reference_state("random_phase", prm=prm)
    F = np.fft.fftn(s.interior) / n ** 3
    x0 = (0.5 / n - 0.5) * L          # the first cell centre
    want = np.zeros_like(F)
    for q in range(20):
        k = prm[5 * q:5 * q + 3].astype(int)
        amp = prm[5 * q + 3] + 1j * prm[5 * q + 4]
        want[tuple(k % n)] += amp * np.exp(2j * np.pi * k.sum() * x0 / L)
    assert np.abs(F - want).max() < 1e-12


def test_random_points_and_dipole_windings():
    """setup_random_points' 20 drawn lines (all positive, as the reference's draw in
    [0.9, 1] makes them) and setup_add_dipole's pair: the plaquette windings of every
    z-plane sum to the lines' charges, each found within one cell of a drawn centre."""
    shape, box = (64, 64, 4), (24.0, 24.0, 4.0)
    prm = refic_draws.random_points(11, box[0], box[1], 64)
    s = oracle.Stepper(oracle.params(shape, box, (1.0, 1.0, 1.0)),
                       np.zeros(shape, np.complex128))
    s.reference_state("random_points", prm=prm)
    v = s.vortices(1e-12, 0)
    h = box[0] / shape[0]
    cx, cy = (v[:, 0] + 1) * h - box[0] / 2, (v[:, 1] + 1) * h - box[1] / 2
    centres = prm.reshape(20, 3)[:, :2]
    d = np.hypot(cx[:, None] - centres[None, :, 0], cy[:, None] - centres[None, :, 1])
    assert np.all(d.min(axis=1) < h)
    for z in range(shape[2]):
        assert v[v[:, 2] == z, 3].sum() == 20
    shape, box = (48, 32, 2), (48.0, 24.0, 2.0)
    s = oracle.Stepper(oracle.params(shape, box, (1.0, 1.0, 1.0)),
                       np.full(shape, 1.0 + 0j))
    s.reference_insert("dipole")
    v = s.vortices(1e-12, 0)
    assert sorted(v[v[:, 2] == 0, 3]) == [-1, 1]


def test_vortex_ring_geometry():
    """setup_vortex_rings (src/initial.f90:202-266) with one drawn ring.  Both factors
    take the same azimuth theta = atan2(tz, ty), so ring_p1*conjg(ring_p2) is real: the
    reference's "ring" is a density dip with no circulation (imaginary part at rounding
    level).  The dip's zeros, d(r2) = 0, lie on the circle of radius ring_rad about the
    drawn centre, perpendicular to the rotated x axis: the 40 lowest-density points are
    within a cell of it."""
    n, L, R = 48, 16.0, 3.0
    prm = refic_draws.vortex_rings(5, 1, L, 64)
    s = oracle.Stepper(oracle.params((n, n, n), L, (1.0, 1.0, 1.0)),
                       np.zeros((n, n, n), np.complex128))
    s.reference_state("vortex_rings", ring_rad=R, prm=prm)
    psi = s.interior
    assert np.abs(psi.imag).max() <= 1e-15 * np.abs(psi).max()
    assert psi.real.min() >= 0
    X, Y, Z = initial.mesh((n, n, n), (L, L, L))
    low = np.argsort(np.abs(psi).reshape(-1))[:40]
    pts = np.stack([X.reshape(-1)[low], Y.reshape(-1)[low], Z.reshape(-1)[low]])
    c = prm[:9].reshape(3, 3)
    rot = np.array([c[0], c[1], [c[2][0], -c[2][1], c[2][2]]])   # temp = rot @ x
    t = rot @ pts - prm[9:12, None]
    h = L / n
    assert np.all(np.abs(t[0]) < h) and np.all(np.abs(np.hypot(t[1], t[2]) - R) < h)


def test_add_noise_range():
    """add_noise (src/initial.f90:741-757): psi += 0.01*(2u-1, 2u'-1) per point."""
    shape = (8, 6, 4)
    n = 8 * 6 * 4
    u = refic_draws.noise(3, n)
    s = oracle.Stepper(oracle.params(shape, 8.0, (1.0, 1.0, 1.0)), np.zeros(shape, np.complex128))
    s.reference_insert("noise", prm=u)
    d = s.interior.reshape(-1)
    assert np.abs(d.real).max() <= 0.01 and np.abs(d.imag).max() <= 0.01
    assert np.allclose(d.real, 0.01 * (2 * u[0::2] - 1), rtol=0, atol=1e-17)


# ---- Psi_info / Psi_norm (src/diagnostic.f90:7-25) ------------------------------------
INFO = ["info_f64_ragged", "info_f32_ragged", "info_f64_ragged_np4", "info_f32_ragged_np4"]


@pytest.mark.parametrize("name", INFO)
def test_psi_info_matches_reference_printout(name):
    """The oracle's max|psi|, min|psi|, mean|psi| and sum|psi|^2 equal the values the
    reference prints every `shots` steps (src/run.f90:59-63), at the print precision of
    list-directed output (17 digits FP64, 9 FP32), on a ragged 12x10x8 grid, run on 1 and
    on 4 ranks (the 4-rank means and sums are reduced over ranks and divided by the
    global point count, src/diagnostic.f90:16-17)."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    prec, shots = int(f["prec"]), int(f["shots"])
    box = tuple(float(b) for b in f["box"])
    harm = tuple(float(h) for h in f["harm_w"])
    s = oracle.Stepper(oracle.params(f["psi0"].shape, box, harm, dt=1e-3), f["psi0"],
                       prec=prec)
    # FP64: print rounding and summation order (960 terms); FP32: the reference sums |psi|
    # in FP32 in its own order, the oracle in long double
    rt_max, rt_sum = (1e-15, 1e-13) if prec == 64 else (1e-7, 2e-6)
    for i, row in enumerate(f["rows"]):
        s.rk3(shots)
        if i == 0:
            # the state of the first shot is the reference's own snapshot, bit for bit
            assert np.array_equal(s.interior, f["psi"])
        info = s.info()
        assert row[1] == pytest.approx(s.t, rel=1e-15 if prec == 64 else 1e-7)
        assert info[0] == pytest.approx(row[2], rel=rt_max)
        assert info[1] == pytest.approx(row[3], rel=rt_max)
        assert info[2] == pytest.approx(row[4], rel=rt_sum)
        assert info[3] == pytest.approx(row[5], rel=rt_sum)


def test_psi_info_brute_force():
    """orc_psi_info against a direct evaluation on a ragged grid with zeros and a wide
    dynamic range: max and min of |psi|, the mean over all Nx*Ny*Nz points and the
    exactly rounded sum of |psi|^2 (math.fsum)."""
    rng = np.random.default_rng(5)
    shape = (7, 5, 9)
    psi = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * \
        np.exp(rng.uniform(-6, 3, shape))
    psi[2, 3, 4] = 0.0
    p = oracle.params(shape, (7.0, 5.0, 9.0), (1.0, 1.0, 1.0))
    info = oracle.Stepper(p, psi).info()
    a = np.abs(psi)
    assert info[0] == pytest.approx(a.max(), rel=3e-16) and info[1] == 0.0
    assert info[2] == pytest.approx(math.fsum(a.ravel()) / a.size, rel=1e-15)
    assert info[3] == pytest.approx(math.fsum((a * a).ravel()), rel=1e-15)
    # planted mistakes the checks above tell apart: the mean over a quarter of the points
    # (a per-rank count) and sum|psi| in place of sum|psi|^2
    assert abs(info[2] - 4 * info[2]) > 1e-3 * info[2]
    assert abs(info[3] - math.fsum(a.ravel())) > 1e-3 * info[3]


def test_imag_rk3_residual_gate():
    """The gated stop rule: with rgate the run continues past the ungated stop until the
    residual of the step's input is at or below the gate, and ends at a lower energy."""
    shape, box, om, g = (16, 16, 16), (10.0, 10.0, 10.0), (1.0, 1.1, 0.9), 20.0
    pr = initial.Problem(shape, box, om, g, 0.3, -5e-3j)
    kp = pr.kernel_params()
    psi0 = initial.tf_profile(shape, box, om, g, norm=1.0)[0] * \
        initial.phase_noise(shape, 5, 0.1)
    mk = lambda: oracle.Stepper(oracle.params(shape, box, kp["harm_w"],  # noqa: E731
                                              omega_z=kp["omega_r"], c0=kp["c0"],
                                              gnl=kp["gnl"], dt=kp["dt"]), psi0)
    h0 = mk().imag_rk3(20000, 1.0, tol=1e-7)
    gate = h0[-1][3] / 4
    h1 = mk().imag_rk3(20000, 1.0, tol=1e-7, rgate=gate)
    assert len(h1) > len(h0) and len(h1) < 20000
    assert h1[-1][3] <= gate and h1[-1][1] < h0[-1][1]
    # the last 8 steps all met both conditions
    for k in range(1, 9):
        assert abs(h1[-k][1] - h1[-k - 1][1]) < 1e-7 * abs(h1[-k][1]) and h1[-k][3] <= gate
