"""GPU parity of the physics beyond the harmonic trap: the reference's potential kinds,
laser beam, switch-off times and advective term (src/finite_diff.f90:30-130), the mobius
Trap table (src/initial.f90:160-196) and the forcing's vortex imprint
(src/force.f90:32-53), through the C ABI, against the reference's own outputs (golden
fixtures) and the CPU oracle on the same inputs.

Tolerances: psi within 1e-12 (FP64) / 2e-6 (FP32) relative max-norm, as for the harmonic
fixtures.  Potential tables: bit for bit for the kinds built from +, -, *, / and sqrt
(correctly rounded on both sides); within 8 ulp of max|pot| where pow, atan2, acos, sin,
cos or exp are evaluated per point (device and host math libraries may differ in the
last place)."""
import glob
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, oracle
from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu

POT = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "pot_*.npz")))
EXACT_KINDS = ["zero", "harmonic", "harmonic_ring", "barrier", "box_trap", "channel",
               "quartic"]
LIBM_KINDS = ["harmonic_ring2", "sphere"]


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def announce(msg):
    print(f"\n[case] {msg}", flush=True)


def _opt(f, key, default):
    if key not in f:
        return default
    v = f[key]
    if v.dtype.kind == "f" and np.isinf(v):
        return None
    return v.item()


def _physics(f):
    return dict(potential=str(f["potential"]), lbeam=bool(f["lbeam"]),
                wbeam=float(f["wbeam"]), potential_off=_opt(f, "potential_off", None),
                t_spindown=_opt(f, "t_spindown", None),
                advective_vx=float(f["advective_vx"]),
                two_dim=bool(f["two_dim"]) if "two_dim" in f else False)


def _condensate(f):
    box = tuple(float(b) for b in f["box"])
    return tbec.Condensate(f["psi0"].shape, box, tuple(float(h) for h in f["harm_w"]),
                           omega_z=float(f["omega_z"]), dt=complex(f["dt"]),
                           prec=int(f["prec"]))


@pytest.mark.parametrize("stored", [False, True])
@pytest.mark.parametrize("name", POT)
def test_potential_fixture_vs_reference(name, stored):
    """20 real-time RK3 steps (mobius: 10) from the fixture's state and time against the
    reference build: time-dependent openings, the box trap's tilt refreshed every step,
    switch-off times crossed mid-run, the laser, the advective term and the mobius Trap
    built on the device."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    prec = int(f["prec"])
    announce(f"{name} shape={f['psi0'].shape} prec={prec} w_stored={stored}")
    with _condensate(f) as c:
        c.set_w_storage(stored)
        c.set_physics(**_physics(f))
        c.upload(f["psi0"])
        c.t = float(f["t0"])
        c.step(int(f["steps"]))
        got = c.download()
        t_end = c.t
    tol = 1e-12 if prec == 64 else 2e-6
    assert relmax(got, f["psi"]) < tol
    assert t_end == pytest.approx(float(f["t"]), rel=1e-6 if prec == 32 else 1e-15)


def _pair(prec, kind, shape=(20, 18, 16), box=(16.0, 14.0, 12.0), t=0.0, **kw):
    harm = tuple(L * L for L in box)
    p = oracle.params(shape, box, harm, potential=kind, **kw)
    psi0 = np.zeros(shape, dtype=np.complex128 if prec == 64 else np.complex64)
    s = oracle.Stepper(p, psi0, prec=prec, t=t)
    c = tbec.Condensate(shape, box, harm, dt=1e-3, prec=prec)
    c.set_physics(kind, **kw)
    return s, c


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("kind", EXACT_KINDS)
def test_potential_table_bitwise(kind, prec):
    """pot from the device equals the oracle's bit for bit (the oracle equals the
    reference bit for bit); at t = 49.99 the openings are still closing, with a tilt."""
    announce(f"potential table {kind} prec={prec}")
    s, c = _pair(prec, kind)
    with c:
        for t in (0.0, 49.99, 60.0):
            got = c.potential(t)
            ref = s.potential(t)
            assert got.dtype == ref.dtype
            assert np.array_equal(got, ref), f"t={t}: {np.sum(got != ref)} points differ"


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("kind", LIBM_KINDS)
def test_potential_table_libm(kind, prec):
    """Kinds that call pow / atan2 / acos / sin / cos per point: same thresholds decided
    (no point flips between the discrete levels) and values within 8 ulp."""
    announce(f"potential table {kind} prec={prec}")
    s, c = _pair(prec, kind, shape=(24, 24, 24), box=(16.0, 16.0, 16.0))
    with c:
        got = c.potential(49.995)
        ref = s.potential(49.995)
    eps = np.finfo(ref.dtype).eps
    assert np.abs(got - ref).max() <= 8 * eps * np.abs(ref).max()


@pytest.mark.parametrize("prec", [64, 32])
def test_laser_and_switch_off(prec):
    """The laser term on a zero potential (exp per point) and the switch-off: pot = 0
    for t > potential_off, compared in the working precision."""
    announce(f"laser prec={prec}")
    s, c = _pair(prec, "zero", shape=(48, 16, 16), box=(160.0, 32.0, 16.0), lbeam=True,
                 wbeam=8.0, potential_off=1.0)
    with c:
        got, ref = c.potential(0.5), s.potential(0.5)
        eps = np.finfo(ref.dtype).eps
        assert ref.max() > 1.0
        assert np.abs(got - ref).max() <= 8 * eps * np.abs(ref).max()
        assert not c.potential(1.0 + 1e-3).any()         # t > potential_off: zero
        assert np.array_equal(c.potential(1.0), got)     # t == potential_off: still on


@pytest.mark.parametrize("prec", [64, 32])
def test_mobius_trap(prec):
    """The mobius Trap built on the device (40000-term sums, capped at 3, max - Trap)
    against the oracle's; 33 points in y put a point at y = 0, so the bump at (25, 0)
    is on the grid."""
    shape, box = (32, 33, 16), (64.0, 64.0, 16.0)
    announce(f"mobius {shape} prec={prec}")
    s, c = _pair(prec, "mobius", shape=shape, box=box)
    s.set_trap(s.mobius_trap())
    with c:
        got = c.potential(0.0)
    ref = s.potential(0.0)
    tol = 1e-12 if prec == 64 else 2e-5
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    assert ref.max() > 10.0 and ref.min() == 0.0


@pytest.mark.parametrize("prec", [64, 32])
def test_table_kind_matches_harmonic(prec):
    """A caller's table holding the harmonic trap steps exactly like the analytic
    harmonic kernel, and like the oracle's table kind."""
    shape, box = (20, 18, 16), (16.0, 14.0, 12.0)
    announce(f"table kind {shape} prec={prec}")
    harm = tuple(L * L for L in box)
    psi0 = initial.tf_profile(shape, box, (1.0, 1.0, 1.0), 500.0, mu=10.0)[0] * \
        math.sqrt(1000.0) + initial.noise(shape, 3, 1e-3)
    cdt = np.complex128 if prec == 64 else np.complex64
    psi0 = psi0.astype(cdt)
    p = oracle.params(shape, box, harm)
    table = oracle.Stepper(p, psi0, prec=prec).potential(0.0)
    with tbec.Condensate(shape, box, harm, dt=1e-3, prec=prec) as c:
        c.upload(psi0)
        c.step(30)
        analytic = c.download()
    with tbec.Condensate(shape, box, harm, dt=1e-3, prec=prec) as c:
        c.set_physics("table")
        with pytest.raises(tbec.TbecError):
            c.step(1)                      # no table given yet
        c.set_potential_table(table)
        c.upload(psi0)
        c.t = 0.0
        c.step(30)
        tab = c.download()
    s = oracle.Stepper(oracle.params(shape, box, harm, potential="table"), psi0, prec=prec)
    s.set_trap(table)
    s.rk3(30)
    tol = 1e-12 if prec == 64 else 2e-6
    assert relmax(tab, analytic) < tol
    assert relmax(tab, s.interior) < tol


def test_switch_times_long_run_with_graphs():
    """64 steps with the potential switched off after step 10 and the frame term after
    step 15: the steps after the switches run as captured graphs (16-step chunks)."""
    shape, box = (16, 16, 16), (16.0, 16.0, 16.0)
    announce("switch times, 64 steps")
    f = np.load(os.path.join(GOLDEN, "pot_harm_switch_f64.npz"))
    harm = tuple(float(h) for h in f["harm_w"])
    kw = dict(potential_off=0.0105, t_spindown=0.0155)
    s = oracle.Stepper(oracle.params(shape, box, harm, omega_z=-1.4, **kw), f["psi0"])
    s.rk3(64)
    with tbec.Condensate(shape, box, harm, omega_z=-1.4, dt=1e-3, prec=64) as c:
        c.set_physics("harmonic", **kw)
        c.upload(f["psi0"])
        for n in (5, 3, 40, 16):          # calls that straddle the switches
            c.step(n)
        got = c.download()
        assert c.t == pytest.approx(s.t, rel=1e-15)
    assert relmax(got, s.interior) < 1e-12


def test_switch_after_replayed_graphs():
    """One 80-step call whose switch-offs fall after 32 and 48 replayed steps: the
    captured graphs of the first run are dropped while their replays may still be in
    flight (the library drains the stream first), and the result matches the oracle."""
    shape, box = (16, 16, 16), (16.0, 16.0, 16.0)
    announce("switch times after replayed graphs, 80 steps")
    f = np.load(os.path.join(GOLDEN, "pot_harm_switch_f64.npz"))
    harm = tuple(float(h) for h in f["harm_w"])
    kw = dict(potential_off=0.0325, t_spindown=0.0485)
    s = oracle.Stepper(oracle.params(shape, box, harm, omega_z=-1.4, **kw), f["psi0"])
    s.rk3(80)
    with tbec.Condensate(shape, box, harm, omega_z=-1.4, dt=1e-3, prec=64) as c:
        c.set_physics("harmonic", **kw)
        c.upload(f["psi0"])
        c.step(80)
        got = c.download()
        assert c.t == pytest.approx(s.t, rel=1e-15)
    assert relmax(got, s.interior) < 1e-12


def test_box_trap_imaginary_time_vs_oracle():
    """Renormalised imaginary-time RK3 in the box trap (tabulated potential, evaluated at
    the handle's time, t = 0 here as in the reference's imaginary loop) against the oracle."""
    shape, box = (24, 24, 24), (16.0, 16.0, 16.0)
    announce("box trap imaginary time")
    harm = tuple(L * L for L in box)
    psi0 = (initial.noise(shape, 9, 1.0) + 1.0).astype(np.complex128)
    p = oracle.params(shape, box, harm, c0=1.0, gnl=1.0, dt=-2e-3j, potential="box_trap")
    s = oracle.Stepper(p, psi0)
    s.imag_rk3(30, 1.0)
    with tbec.Condensate(shape, box, harm, dt=-2e-3j, prec=64) as c:
        c.set_physics("box_trap")
        c.upload(psi0)
        c.step(30, tbec.IMAG_RK3)
        got = c.download()
    assert relmax(got, s.interior) < 1e-12


@pytest.mark.parametrize("prec", [64, 32])
def test_imprint_vortex_vs_oracle(prec):
    """force_random_points' imprint (pade(r) exp(i dir theta)) against the oracle, twice
    with opposite senses."""
    shape, box = (24, 20, 16), (16.0, 14.0, 12.0)
    announce(f"imprint prec={prec}")
    harm = tuple(L * L for L in box)
    cdt = np.complex128 if prec == 64 else np.complex64
    psi0 = (initial.noise(shape, 4, 0.1) + 1.0).astype(cdt)
    s = oracle.Stepper(oracle.params(shape, box, harm), psi0, prec=prec)
    s.imprint_vortex(1.3, -2.1, 1.0)
    s.imprint_vortex(-3.2, 0.7, -1.0)
    with tbec.Condensate(shape, box, harm, dt=1e-3, prec=prec) as c:
        c.upload(psi0)
        c.imprint_vortex(1.3, -2.1, 1.0)
        c.imprint_vortex(-3.2, 0.7, -1.0)
        got = c.download()
        with pytest.raises(tbec.TbecError):
            c.imprint_vortex(0.0, 0.0, 0.5)
    tol = 1e-13 if prec == 64 else 2e-6
    assert relmax(got, s.interior) < tol
