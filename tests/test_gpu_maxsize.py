"""Large grids: 64-bit indexing past 2^32 padded elements, against a small run.

With periodic boundaries on every face, no trap (harm_w = 0) and no rotation, one stage
computes the same arithmetic at every point from the same neighbours.  So a 1664^3
grid (4.6e9 points; each padded buffer holds 4.67e9 elements, 37 GB) filled with 26^3
copies of a periodic 64^3 state must hold, after the same steps, 26^3 copies of the
64^3 run's result, bit for bit.  Both handles use the same stage kernel (tile 2, FP32,
centred sums, w rebuilt); the 64^3 run is itself pinned against the oracle by the
parity tests."""
import numpy as np
import pytest

from trapbec_amd import tbec

pytestmark = pytest.mark.gpu


def periodic_state(n, seed=11):
    """A smooth periodic field plus seeded noise on an n^3 grid (complex64)."""
    x = 2.0 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    psi = (1.0 + 0.3 * np.sin(X) * np.cos(2 * Y) + 0.2j * np.cos(Z + 0.5)
           + 0.1 * np.exp(1j * (X - 2 * Z)))
    rng = np.random.default_rng(seed)
    psi = psi + 1e-3 * (rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n)))
    return psi.astype(np.complex64)


def run(psi, n, h, steps):
    with tbec.Condensate((n, n, n), (n * h,) * 3, (0.0, 0.0, 0.0), gnl=1.0, dt=1e-3,
                         prec=32, tile=2) as c:
        c.set_boundaries("periodic", True)
        c.upload(psi)
        c.step(steps)
        return c.download()


@pytest.mark.timeout(600)
def test_periodic_tiling_past_2_to_the_32_elements():
    n0, reps, h, steps = 64, 26, 0.25, 4
    n1 = n0 * reps
    print(f"\n[case] periodic {n0}^3 tiled {reps}^3 times = {n1}^3 FP32, {steps} steps",
          flush=True)
    small0 = periodic_state(n0)
    small = run(small0, n0, h, steps)
    assert not np.array_equal(small, small0)          # the steps changed the state
    big = np.tile(small0, (reps, reps, reps))
    print(f"[case] {n1}^3 state built ({big.nbytes / 2**30:.1f} GiB)", flush=True)
    got = run(big, n1, h, steps)
    del big
    print("[case] large run done; comparing tiles", flush=True)
    v = got.reshape(reps, n0, reps, n0, reps, n0)
    for a in range(reps):       # one x-slab of tiles at a time keeps the temporaries small
        assert np.array_equal(v[a], np.broadcast_to(small[:, None, :, None, :],
                                                     (n0, reps, n0, reps, n0))), a
