"""Vortex-core detection on the GPU against the oracle's plaquette winding."""
import numpy as np
import pytest

from conftest import oracle
from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [64, 32])
def test_vortex_cores_match_oracle(prec):
    shape, box = (48, 40, 36), (12.0, 10.0, 9.0)
    print(f"\n[case] vortex cores {shape} prec={prec}", flush=True)
    f = initial.tf_profile(shape, box, (0.4, 0.4, 0.4), 1.0, mu=10.0)[0]
    f = f * initial.vortex_ring_factor(shape, box, 2.5, 0.3, 0.4)
    for x0, y0, q in ((-3.1, 2.2, 1), (2.6, -1.3, -1), (0.4, 3.3, 2)):
        f = f * initial.vortex_line_factor(shape, box, x0, y0, q, 0.4)
    f = f + initial.noise(shape, 3, 1e-3)
    if prec == 32:
        f = f.astype(np.complex64).astype(np.complex128)
    s = oracle.Stepper(oracle.params(shape, box, (0.0, 0.0, 0.0)), f)
    with tbec.Condensate(shape, box, (0.0, 0.0, 0.0), prec=prec) as c:
        c.upload(f)
        for o in (0, 1, 2):
            got = c.vortex_cores(1e-4, o)
            ref = s.vortices(1e-4, o)
            assert len(ref) > 0
            assert np.array_equal(got, ref), (o, len(got), len(ref))
