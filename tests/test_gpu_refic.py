"""GPU parity of the reference's own initial states and insertions
(src/initial.f90:58-156,202-870; src/run.f90:26,37-40), generated on the device through
the C ABI: against the reference's snapshots (golden fixtures) and, for the states the
reference draws from its clock-seeded generator, against the oracle fed the same draws.

Tolerance: relative max-norm 1e-12 (FP64) / 2e-5 (FP32).  Each point is a product of up
to 100 vortex factors whose pow / exp / atan2 / cos / sin come from the device library,
each within a few ulp of the host's."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, oracle
import refic_draws
from trapbec_amd import tbec

pytestmark = pytest.mark.gpu

FIXTURES = sorted(os.path.basename(p)[:-4] for p in
                  glob.glob(os.path.join(GOLDEN, "ic_*.npz")) +
                  glob.glob(os.path.join(GOLDEN, "ins_*.npz")))


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def announce(msg):
    print(f"\n[case] {msg}", flush=True)


def tol(prec):
    return 1e-12 if prec == 64 else 2e-5


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_state_vs_fixture(name):
    """The state, then the insertions in the reference's order (phase_imprint in the
    imaginary stage, then add_pcurrent, add_ring, add_dipole)."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    prec = int(f["prec"])
    shape = f["psi"].shape
    box = tuple(float(b) for b in f["box"])
    announce(f"{name} shape={shape} prec={prec}")
    wn, rr = int(f["winding_number"]), float(f["ring_rad"])
    with tbec.Condensate(shape, box, tuple(float(h) for h in f["harm_w"]), dt=0.0,
                         prec=prec) as c:
        c.set_physics(str(f["potential"]))
        c.reference_state(str(f["init_cond"]), winding_number=wn, ring_rad=rr)
        if bool(f["phase_imprint_on"]):
            c.reference_insert("phase_imprint", winding_number=wn)
        if bool(f["add_pcurrent"]):
            c.reference_insert("pcurrent")
        if bool(f["add_ring"]):
            c.reference_insert("ring", ring_rad=rr)
        if bool(f["add_dipole"]):
            c.reference_insert("dipole")
        got = c.download()
    assert relmax(got, f["psi"]) < tol(prec)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("kind", ["random_phase", "random_points", "vortex_rings"])
def test_drawn_states_vs_oracle(kind, prec):
    """States whose values the reference draws at random: the device's seeded draws,
    replicated for the oracle (refic_draws), give the oracle's state."""
    shape, box, seed = (24, 20, 16), (16.0, 14.0, 12.0), 99
    announce(f"{kind} prec={prec}")
    harm = (4.0, 4.0, 4.0)
    if kind == "random_phase":
        prm = refic_draws.random_phase(seed, prec)
    elif kind == "random_points":
        prm = refic_draws.random_points(seed, box[0], box[1], prec)
    else:
        prm = refic_draws.vortex_rings(seed, 3, box[0], prec)
    cdt = np.complex128 if prec == 64 else np.complex64
    s = oracle.Stepper(oracle.params(shape, box, harm), np.zeros(shape, cdt), prec=prec)
    s.reference_state(kind, ring_rad=2.5, prm=prm)
    with tbec.Condensate(shape, box, harm, dt=0.0, prec=prec) as c:
        c.reference_state(kind, ring_rad=2.5, vor_ring_count=3, seed=seed)
        got = c.download()
    assert relmax(got, s.interior) < tol(prec)


@pytest.mark.parametrize("prec", [64, 32])
def test_noise_insert_vs_oracle(prec):
    """add_noise with the seeded per-point draws, on top of the constant state."""
    shape, box, seed = (24, 20, 16), (16.0, 14.0, 12.0), 5
    announce(f"noise prec={prec}")
    cdt = np.complex128 if prec == 64 else np.complex64
    s = oracle.Stepper(oracle.params(shape, box, (1.0, 1.0, 1.0)), np.zeros(shape, cdt),
                       prec=prec)
    s.reference_state("constant")
    s.reference_insert("noise", prm=refic_draws.noise(seed, int(np.prod(shape))))
    with tbec.Condensate(shape, box, (1.0, 1.0, 1.0), dt=0.0, prec=prec) as c:
        c.reference_state("constant")
        c.reference_insert("noise", seed=seed)
        got = c.download()
        with pytest.raises(tbec.TbecError):
            c.reference_state("vortex_rings", vor_ring_count=9)
    assert np.array_equal(got, s.interior)
