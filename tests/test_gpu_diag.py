"""GPU parity of the output-time diagnostics (velocity from the phase, quantum
pressure) against the reference's own vel/qpress records and the oracle."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, oracle
from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name", ["diag_f64_16", "diag_f32_16"])
def test_diagnostics_vs_reference_fixture(name):
    """libm-level differences of atan2/hypot only: FP64 within 1e-12, FP32 within 2e-5
    relative max-norm of the reference's fields."""
    f = np.load(os.path.join(GOLDEN, name + ".npz"))
    prec = int(f["prec"])
    print(f"\n[case] {name}", flush=True)
    box = tuple(float(b) for b in f["box"])
    with tbec.Condensate(f["psi"].shape, box, tuple(L * L for L in box), prec=prec) as c:
        c.upload(f["psi"])
        v = c.velocity()
        q = c.qpressure()
        back = c.download()
    tol = 1e-12 if prec == 64 else 2e-5
    assert relmax(v, f["vel"]) < tol
    assert relmax(q, f["qpress"]) < tol
    assert np.array_equal(back, f["psi"])        # the state is untouched


@pytest.mark.parametrize("prec", [64, 32])
def test_diagnostics_ragged_after_imaginary_steps(prec):
    """A ragged grid after renormalised imaginary-time steps (pending scale applied
    first) against the oracle on the downloaded state."""
    shape, box = (37, 23, 50), (9.0, 7.0, 11.0)
    print(f"\n[case] diagnostics ragged {shape} prec {prec}", flush=True)
    pr = initial.Problem(shape, box, (1.0, 1.2, 0.8), 50.0, 0.3, -2e-3j)
    psi0 = initial.tf_profile(shape, box, pr.omega, pr.g, norm=1.0)[0] * \
        initial.vortex_line_factor(shape, box, 0.4, -0.3, 1, 0.4) * \
        initial.phase_noise(shape, 3, 0.2)
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        c.upload(psi0)
        c.step(6, tbec.IMAG_RK3)
        v = c.velocity()
        q = c.qpressure()
        psi = c.download()
    s = oracle.Stepper(oracle.params(shape, box, (1.0, 1.0, 1.0)), psi, prec=prec)
    tol = 1e-12 if prec == 64 else 2e-5
    assert relmax(v, s.velocity()) < tol
    assert relmax(q, s.qpressure()) < tol
