"""Batched ensembles (many condensates per launch) against one-at-a-time solves."""
import numpy as np
import pytest

from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu


def _problems():
    grids = ((24, 24, 24), (32, 20, 28), (16, 16, 40))
    return initial.config_c4(seed=7, n_inst=5, dx=0.4, dtau=4e-3, grids=grids)


@pytest.mark.parametrize("prec", [64, 32])
def test_batch_equals_single_solves(prec):
    """Each instance of a batch solve ends bitwise equal to its own single-handle solve
    (same kernels, same tiles, same fixed-order reductions), with the same step count."""
    probs = _problems()
    print(f"\n[case] batch of {len(probs)} prec={prec}", flush=True)
    with tbec.Batch([p for p, _ in probs], prec=prec) as b:
        for i, (_, psi0) in enumerate(probs):
            b.upload(i, psi0)
        res = b.imag_solve(2500, 1e-9, 1.0, 40)
        got = [b.download(i) for i in range(len(probs))]
    # the instances finish at different checks (830 to 2049 steps in the FP64 oracle), so the launches drop finished instances'
    # work items along the way
    assert len({int(r["steps"]) for r in res}) > 1
    for i, (pr, psi0) in enumerate(probs):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.upload(psi0)
            r = c.imag_solve(2500, 1e-9, 1.0, 40)
            ref = c.download()
        assert r["steps"] == res[i]["steps"], i
        assert r["converged"] == res[i]["converged"]
        assert np.array_equal(got[i], ref), (i, np.abs(got[i] - ref).max())
        assert res[i]["energy"] == r["energy"]


@pytest.mark.parametrize("prec", [64, 32])
def test_batch_second_solve_repeats_the_first(prec):
    """A solve drops finished instances from its launches; the next solve starts with
    every instance again: re-uploading the starts repeats the first solve bit for bit."""
    probs = _problems()
    print(f"\n[case] batch solved twice prec={prec}", flush=True)
    with tbec.Batch([p for p, _ in probs], prec=prec) as b:
        runs = []
        for _ in range(2):
            for i, (_, psi0) in enumerate(probs):
                b.upload(i, psi0)
            res = b.imag_solve(2500, 1e-9, 1.0, 40)
            runs.append((res, [b.download(i) for i in range(len(probs))]))
    (r1, g1), (r2, g2) = runs
    assert [int(r["steps"]) for r in r1] == [int(r["steps"]) for r in r2]
    for i in range(len(probs)):
        assert np.array_equal(g1[i], g2[i]), i


@pytest.mark.parametrize("prec", [64, 32])
def test_batch_upload_after_solve_round_trips(prec):
    """An upload after a solve replaces the instance's state as given: the download
    returns it bit for bit (no renormalisation scale left over from the solve)."""
    probs = _problems()[:3]
    print(f"\n[case] batch re-upload after solve prec={prec}", flush=True)
    with tbec.Batch([p for p, _ in probs], prec=prec) as b:
        for i, (_, psi0) in enumerate(probs):
            b.upload(i, psi0)
        b.imag_solve(100, 1e-30, 1.0, 20)
        x = (2.0 * probs[1][1]).astype(b.cdt)
        b.upload(1, x)
        got = b.download(1)
    assert np.array_equal(got, x)


def test_batch_rejects_mixed_precision():
    probs = _problems()[:2]
    with pytest.raises(tbec.TbecError):
        tbec.Batch([p for p, _ in probs], prec=48)
