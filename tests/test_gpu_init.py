"""Seeded initial states generated on the device equal the numpy generator's."""
import numpy as np
import pytest

from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n", [("m", 48), ("c1", 32), ("c2", 40), ("c3", 36)])
@pytest.mark.parametrize("prec", [64, 32])
def test_device_initial_state_matches_numpy(name, n, prec):
    print(f"\n[case] device initial state {name} n={n} prec={prec}", flush=True)
    pr, spec = initial.ic_spec(name, n)
    ref = {"m": lambda: initial.config_m(n)[1], "c1": lambda: initial.config_c1(n=n)[1],
           "c2": lambda: initial.config_c2(n=n)[1], "c3": lambda: initial.config_c3(n=n)[1]}[name]()
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        c.init_state(spec)
        got = c.download()
    tol = 1e-12 if prec == 64 else 2e-7
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
