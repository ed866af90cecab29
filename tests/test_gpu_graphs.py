"""Captured graphs against eager launches, bit for bit.

Real time: tbec_step runs whole 16-step chunks as graph replays (captured on first use)
and the remainder eagerly.  One handle advances by step(16) several times in a row, with
a host copy (a synchronisation) between the replays.  A twin handle advances by sixteen
step(1) calls, which never capture.

Imaginary time: each tbec_step call ends by folding the pending renormalisation scale
into psi, so the comparison keeps the call structure fixed.  The imaginary-time solve
replays whole chunks of check_every steps (rounded up to even) as graphs and launches a
shorter remainder eagerly.  Each solve call covers 16 steps: with check_every 16 that is
one graph replay (the graph persists across calls), with check_every 17 (a chunk of 18)
the 16 steps are launched eagerly, with the same single convergence check at the end.
The two handles must agree after every call in psi, energy, mu, residual and norm, with
reductions, renormalisation and the finalisers inside the graphs."""
import numpy as np
import pytest

from trapbec_amd import initial, tbec

pytestmark = pytest.mark.gpu

SHAPE, BOX, OM, G = (40, 36, 32), (14.0, 12.0, 11.0), (1.0, 1.2, 0.9), 50.0


def start():
    return (initial.tf_profile(SHAPE, BOX, OM, G, norm=1.0)[0] *
            initial.phase_noise(SHAPE, 3, 0.2)).astype(np.complex128)


@pytest.mark.parametrize("prec", [64, 32])
def test_real_time_graph_replays_equal_eager_steps(prec):
    print(f"\n[case] graph vs eager real prec={prec} {SHAPE}", flush=True)
    psi0 = start()
    pr = initial.Problem(SHAPE, BOX, OM, G, 0.4, 1e-3)
    with tbec.Condensate.from_problem(pr, prec=prec) as a, \
            tbec.Condensate.from_problem(pr, prec=prec) as b:
        a.upload(psi0)
        b.upload(psi0)
        for rep in range(4):
            a.step(16)                        # graph: captured at rep 0, replayed after
            for _ in range(16):
                b.step(1)                     # eager
            ga, gb = a.download(), b.download()
            assert np.array_equal(ga, gb), f"replay {rep}"
        assert not np.array_equal(ga, psi0.astype(ga.dtype))


@pytest.mark.parametrize("prec", [64, 32])
def test_imaginary_time_graph_replays_equal_eager_steps(prec):
    print(f"\n[case] graph vs eager imaginary time prec={prec} {SHAPE}", flush=True)
    psi0 = start()
    pr = initial.Problem(SHAPE, BOX, OM, G, 0.4, -1e-3j)
    with tbec.Condensate.from_problem(pr, prec=prec) as a, \
            tbec.Condensate.from_problem(pr, prec=prec) as b:
        a.upload(psi0)
        b.upload(psi0)
        for rep in range(4):
            ra = a.imag_solve(16, 1e-30, 1.0, 16)   # one graph replay per call
            rb = b.imag_solve(16, 1e-30, 1.0, 17)   # the same sixteen steps, eager
            assert ra["steps"] == rb["steps"] == 16
            for k in ("energy", "mu", "residual", "norm", "rel_change"):
                assert ra[k] == rb[k], (rep, k)
            ga, gb = a.download(), b.download()
            assert np.array_equal(ga, gb), f"replay {rep}"
        assert not np.array_equal(ga, psi0.astype(ga.dtype))
