"""tbec_diagnose_noalloc (reduction partials in a static device array) against
tbec_diagnose (partials in the handle's own buffer), through the C ABI: equal bit for bit
on two calls in a row, the second on a smaller grid, and a clear rejection of a
configuration needing more partial slots than the static array holds."""
import numpy as np
import pytest

from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu

FIELDS = ("norm", "energy", "mu", "residual", "max_abs", "min_abs", "mean_abs", "int_psi2")


def _state(shape, box, om, g):
    psi = initial.tf_profile(shape, box, om, g, norm=1.0)[0] * \
        initial.vortex_line_factor(shape, box, 0.4, -0.3, 1, 0.3)
    return (psi + initial.noise(shape, 3, 1e-3)).astype(np.complex128)


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_diagnose_equals_diagnose(prec):
    cases = [((48, 40, 36), (12.0, 10.0, 9.0), 0.3),   # the larger grid first
             ((20, 16, 12), (10.0, 8.0, 6.0), 0.0)]    # then a smaller one
    for shape, box, Omega in cases:
        print(f"\n[case] diagnose_noalloc prec={prec} shape={shape} Omega={Omega}", flush=True)
        om, g = (1.0, 1.1, 0.9), 40.0
        pr = initial.Problem(shape, box, om, g, Omega, 1e-3)
        psi = _state(shape, box, om, g)
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.upload(psi)
            got = c.diagnose_noalloc()      # first: before the handle has any buffer
            ref = c.diagnose()
            again = c.diagnose_noalloc()    # after the handle's own buffer exists
        for k in FIELDS:
            assert got[k] == ref[k], (k, got[k], ref[k])
            assert again[k] == ref[k], (k, again[k], ref[k])
        assert np.isfinite(ref["energy"]) and ref["norm"] > 0


def test_noalloc_diagnose_rejects_too_many_slots():
    """64 tiles per plane with one-plane segments over 520 planes need 33280 slots."""
    shape = (520, 256, 256)
    print(f"\n[case] diagnose_noalloc rejection shape={shape}", flush=True)
    pr = initial.Problem(shape, (26.0, 12.8, 12.8), (1.0, 1.0, 1.0), 10.0, 0.0, 1e-3)
    with tbec.Condensate.from_problem(pr, prec=32) as c:
        c.set_tuning(2, 1)
        with pytest.raises(tbec.TbecError) as e:
            c.diagnose_noalloc()
        assert e.value.code == -1
        assert "reduction slots" in str(e.value)
