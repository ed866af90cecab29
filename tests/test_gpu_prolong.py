"""tbec_prolong (trilinear interpolation of one handle's psi onto another grid of the same
box) and the coarse-to-fine ground-state solve built on it.

The interpolation is checked against scipy's linear map_coordinates (edge values
replicated, which is the kernel's clamping) on anisotropic grids, finer and coarser; the
same grid must copy psi bit for bit; the source's pending renormalisation must be
applied; mismatched handles are refused.  The continuation solve must land on the ground
state the one-grid solve finds."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from trapbec_amd import initial, tbec
from trapbec_amd import warmstart as ws

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[case] {msg}", flush=True)


def handle(shape, box=(6.0, 7.0, 8.0), prec=64, dt=1e-3):
    return tbec.Condensate(shape, box, (1.0, 1.2, 0.8), c0=0.0, gnl=2.0, dt=dt, prec=prec)


def linear(psi, shape):
    """psi sampled at the cell centres of `shape` on the same box, linear, edges held."""
    idx = np.indices(shape).reshape(3, -1).astype(np.float64)
    pos = np.stack([(idx[d] + 0.5) * psi.shape[d] / shape[d] - 0.5 for d in range(3)])
    re = map_coordinates(psi.real.astype(np.float64), pos, order=1, mode="nearest")
    im = map_coordinates(psi.imag.astype(np.float64), pos, order=1, mode="nearest")
    return (re + 1j * im).reshape(shape)


def random_psi(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("prec,tol", [(64, 1e-13), (32, 2e-6)])
@pytest.mark.parametrize("src_shape,dst_shape", [((10, 12, 14), (25, 19, 40)),
                                                 ((24, 20, 18), (9, 13, 7))])
def test_prolong_is_linear_interpolation(prec, tol, src_shape, dst_shape):
    announce(f"prolong fp{prec} {src_shape} -> {dst_shape}")
    psi = random_psi(src_shape, 5)
    with handle(src_shape, prec=prec) as s, handle(dst_shape, prec=prec) as d:
        s.upload(psi)
        tbec.prolong(d, s)
        got = d.download()
    want = linear(psi.astype(s.cdt), dst_shape)
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"rel max {err:.2e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("prec", [64, 32])
def test_prolong_same_grid_copies_bitwise(prec):
    shape = (11, 9, 13)
    announce(f"prolong fp{prec} identity {shape}")
    psi = random_psi(shape, 6)
    with handle(shape, prec=prec) as s, handle(shape, prec=prec) as d:
        s.upload(psi)
        tbec.prolong(d, s)
        assert np.array_equal(d.download(), s.download())


def test_prolong_applies_the_source_renormalisation():
    """After imaginary-time steps the source stores psi unnormalised with a pending scale;
    the prolongation reads it as the download does."""
    src_shape, dst_shape = (16, 18, 20), (30, 28, 34)
    announce(f"prolong after imaginary steps {src_shape} -> {dst_shape}")
    pr = initial.Problem(src_shape, (8.0, 8.0, 8.0), (1.0, 1.0, 1.0), 50.0, 0.0, -2e-3j)
    psi0 = initial.tf_profile(src_shape, pr.box, pr.omega, pr.g, norm=1.0)[0] * \
        initial.phase_noise(src_shape, 3, 0.1)
    with tbec.Condensate.from_problem(pr, prec=64) as s, \
            handle(dst_shape, box=pr.box, prec=64) as d:
        s.upload(psi0 * 3.0)
        s.step(7, tbec.IMAG_RK3)
        tbec.prolong(d, s)
        got = d.download()
        want = linear(s.download(), dst_shape)
    err = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"rel max {err:.2e}", flush=True)
    assert err < 1e-13


def test_prolong_refuses_mismatched_handles():
    announce("prolong argument checks")
    with handle((8, 8, 8), prec=64) as a, handle((12, 12, 12), prec=32) as b, \
            handle((12, 12, 12), box=(6.0, 7.0, 9.0), prec=64) as c:
        for dst, src in ((b, a), (c, a), (a, a)):
            with pytest.raises(tbec.TbecError) as e:
                tbec.prolong(dst, src)
            assert e.value.code == -1


def test_continuation_reaches_the_one_grid_ground_state():
    """A non-rotating FP32 ground state (unique up to a global phase) by the one-grid
    gated solve and by 32^3 -> 64^3 continuation: energies within 2e-7 relative, |psi|
    within 2e-4 relative L2 (each state is within about residual / gap of the ground
    state, the gap being of order the trap frequency), both residuals at or below the
    gate."""
    shape, dtau, gate = (64, 64, 64), 2e-3, 5e-5
    pr = initial.Problem(shape, (16.0, 16.0, 16.0), (1.0, 1.0, 1.2), 300.0, 0.0, -dtau * 1j)
    announce(f"continuation {shape} vs one grid")
    with tbec.Condensate.from_problem(pr, prec=32) as c:
        c.upload(ws.tf_start(pr, 1.0, seed=0))
        c.set_residual_gate(gate)
        r = c.imag_solve(100000, 1e-10, 1.0, 100)
        d = c.diagnose()
        plain = np.abs(c.download())
    cont, h = ws.continuation_solve(pr, factors=(2,), gate=gate, coarse_gate=gate,
                                    max_steps=100000, check_every=100, keep=True,
                                    log=lambda m: print(m, flush=True))
    try:
        lev = np.abs(h.download())
    finally:
        h.close()
    print(f"one grid: {int(r['steps'])} steps E {d['energy']:.10f} res {d['residual']:.2e}; "
          f"continuation levels {[(lv['shape'][0], lv['steps']) for lv in cont['levels']]} "
          f"E {cont['energy']:.10f} res {cont['residual']:.2e}", flush=True)
    assert r["converged"] and cont["converged"]
    assert d["residual"] <= gate and cont["residual"] <= gate
    assert abs(cont["energy"] - d["energy"]) <= 2e-7 * abs(d["energy"])
    assert np.linalg.norm(lev - plain) / np.linalg.norm(plain) < 2e-4


def test_batch_prolong_equals_single_prolong():
    """Every instance of tbec_batch_prolong equals tbec_prolong of the same state bit for
    bit (grids and boxes differ per instance), and the result starts a batch solve."""
    fine = [initial.Problem((20, 16, 24), (6.0, 5.0, 7.0), (1.0, 1.1, 0.9), 80.0, 0.3, -1e-3j),
            initial.Problem((12, 18, 10), (4.0, 6.0, 5.0), (0.9, 0.9, 1.2), 40.0, 0.0, -2e-3j)]
    coarse = [initial.Problem((9, 11, 7), p.box, p.omega, p.g, p.Omega, p.dt) for p in fine]
    announce("batch prolong, 2 instances")
    psis = [random_psi(p.shape, 7 + i) for i, p in enumerate(coarse)]
    with tbec.Batch(coarse, prec=32) as bc, tbec.Batch(fine, prec=32) as bf:
        for i, psi in enumerate(psis):
            bc.upload(i, psi)
        tbec.batch_prolong(bf, bc)
        got = [bf.download(i) for i in range(2)]
        res = bf.imag_solve(20, 1e-12, 1.0, 10)
        assert all(r["steps"] == 20 for r in res)
    for i, (pc, pf) in enumerate(zip(coarse, fine)):
        with tbec.Condensate.from_problem(pc, prec=32) as s, \
                tbec.Condensate.from_problem(pf, prec=32) as d:
            s.upload(psis[i])
            tbec.prolong(d, s)
            assert np.array_equal(got[i], d.download()), i


def test_batch_prolong_refuses_mismatched_batches():
    announce("batch prolong argument checks")
    p = initial.Problem((8, 8, 8), (4.0, 4.0, 4.0), (1.0, 1.0, 1.0), 10.0, 0.0, -1e-3j)
    q = initial.Problem((8, 8, 8), (4.0, 4.0, 5.0), (1.0, 1.0, 1.0), 10.0, 0.0, -1e-3j)
    with tbec.Batch([p, p], prec=32) as a, tbec.Batch([p], prec=32) as b, \
            tbec.Batch([p, q], prec=32) as c, tbec.Batch([p, p], prec=64) as d:
        for dst, src in ((b, a), (c, a), (d, a), (a, a)):
            with pytest.raises(tbec.TbecError) as e:
                tbec.batch_prolong(dst, src)
            assert e.value.code == -1
