"""Error paths through the C ABI: a non-finite psi is reported as TBEC_ERR_NONFINITE by
the FP64 diagnostics, the imaginary-time solve, step-limited imaginary-time steps and the
batch solve, and the handle works again after a finite upload.  NaN is ordinary data for the kernels (nothing faults); the
host reads the FP64 reductions and turns a non-finite norm or energy into the error."""
import numpy as np
import pytest

from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu

ERR_NONFINITE = -4
SHAPE, BOX, OM, G = (24, 20, 16), (12.0, 10.0, 8.0), (1.0, 1.1, 0.9), 20.0


def states():
    good = initial.tf_profile(SHAPE, BOX, OM, G, norm=1.0)[0].astype(np.complex128)
    bad = good.copy()
    bad[5, 7, 3] = np.nan
    return good, bad


@pytest.mark.parametrize("prec", [64, 32])
def test_diagnose_reports_nonfinite(prec):
    print(f"\n[case] diagnose NaN prec={prec}", flush=True)
    good, bad = states()
    with tbec.Condensate.from_problem(initial.Problem(SHAPE, BOX, OM, G, 0.0, 1e-3),
                                      prec=prec) as c:
        c.upload(bad)
        with pytest.raises(tbec.TbecError) as e:
            c.diagnose()
        assert e.value.code == ERR_NONFINITE
        c.upload(good)
        d = c.diagnose()
        assert np.isfinite(d["energy"]) and d["norm"] == pytest.approx(1.0, rel=1e-5)


@pytest.mark.parametrize("prec", [64, 32])
def test_imag_solve_reports_nonfinite_and_recovers(prec):
    print(f"\n[case] imag_solve NaN prec={prec}", flush=True)
    good, bad = states()
    pr = initial.Problem(SHAPE, BOX, OM, G, 0.3, -2e-3j)
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        c.upload(bad)
        with pytest.raises(tbec.TbecError) as e:
            c.imag_solve(40, 1e-10, 1.0, 10)
        assert e.value.code == ERR_NONFINITE
        c.upload(good)
        r = c.imag_solve(40, 1e-30, 1.0, 10)
        assert r["steps"] == 40 and not r["converged"]
        assert np.isfinite(r["energy"]) and r["norm"] == pytest.approx(1.0, rel=1e-6)


def test_batch_solve_reports_nonfinite():
    print("\n[case] batch imag_solve NaN", flush=True)
    good, bad = states()
    probs = [initial.Problem(SHAPE, BOX, OM, G, 0.3, -2e-3j) for _ in range(2)]
    with tbec.Batch(probs, prec=32) as b:
        b.upload(0, good)
        b.upload(1, bad)
        with pytest.raises(tbec.TbecError) as e:
            b.imag_solve(40, 1e-10, 1.0, 10)
        assert e.value.code == ERR_NONFINITE


@pytest.mark.parametrize("prec", [64, 32])
def test_imag_steps_report_nonfinite_and_recover(prec):
    """tbec_step(TBEC_IMAG_RK3) reads the finalisers' flag: a NaN norm is an error, not a
    silently frozen psi (40 steps: two replayed 16-step graphs and 8 eager steps)."""
    print(f"\n[case] imag steps NaN prec={prec}", flush=True)
    good, bad = states()
    pr = initial.Problem(SHAPE, BOX, OM, G, 0.3, -2e-3j)
    with tbec.Condensate.from_problem(pr, prec=prec) as c:
        c.upload(bad)
        with pytest.raises(tbec.TbecError) as e:
            c.step(40, tbec.IMAG_RK3)
        assert e.value.code == ERR_NONFINITE
        c.upload(good)
        c.step(40, tbec.IMAG_RK3)
        d = c.diagnose()
        assert np.isfinite(d["energy"]) and d["norm"] == pytest.approx(1.0, rel=1e-5)
