"""The workload configurations at their configured shapes, bounded in steps.

C2: 256^3 real time, 2000 steps, FP64 and FP32 from one seeded state (ring + 4 lines).
C3: 256^3 rotating imaginary time (Omega = 0.7), a bounded run in FP64 and FP32, and its
    32^3 twin (same physics and start) against the FP64 oracle.
C4: the 48-instance ensemble at its drawn grids, batched against one-at-a-time solves,
    and 3 heterogeneous instances against the FP64 oracle.
C5: the production-size warm-start network (5 -> 4x512 -> 3*32^3) forward and gradients
    against the oracle, and the 32^3 -> 384^3 trilinear upsample.

Tolerances are written in each test; every case prints its name and shape first, and
its measured errors after."""
import math

import numpy as np
import pytest

from conftest import oracle
import mlp_oracle as mo
from trapbec_amd import initial, tbec
from trapbec_amd.configs import _match_displacement

pytestmark = pytest.mark.gpu


def announce(msg):
    print(f"\n[case] {msg}", flush=True)


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---- C2 -------------------------------------------------------------------------------
def test_c2_fp32_tracks_fp64_256():
    """psi relative L2 of FP32 against FP64 below 1e-5 after 2000 steps; norm and energy
    drift equal in both precisions (the drift is the scheme's, not rounding); the same
    z-normal vortex cores, displaced by at most one grid spacing."""
    n, steps = 256, 2000
    pr, spec = initial.ic_spec("c2", n)
    announce(f"C2 {n}^3 real time, {steps} steps, FP64 and FP32")
    res = {}
    for prec in (64, 32):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.init_state(spec)
            d0 = c.diagnose()
            c.step(steps)
            d1 = c.diagnose()
            psi = c.download().astype(np.complex128)
            cores = c.vortex_cores(1e-5 * d0["max_abs"] ** 2, 0)
        res[prec] = dict(psi=psi, cores=cores,
                         norm_drift=(d1["norm"] - d0["norm"]) / d0["norm"],
                         energy_drift=(d1["energy"] - d0["energy"]) / abs(d0["energy"]))
        print(f"fp{prec}: norm drift {res[prec]['norm_drift']:.4e} energy drift "
              f"{res[prec]['energy_drift']:.4e} cores {len(cores)}", flush=True)
    err = rel_l2(res[32]["psi"], res[64]["psi"])
    h = pr.box[0] / n
    mean_d, max_d = _match_displacement(res[32]["cores"], res[64]["cores"], h)
    print(f"psi rel L2 fp32 vs fp64 {err:.3e}; core displacement mean {mean_d} max {max_d} "
          f"(h = {h})", flush=True)
    assert err < 1e-5
    assert abs(res[32]["norm_drift"] - res[64]["norm_drift"]) < 1e-6
    assert abs(res[32]["energy_drift"] - res[64]["energy_drift"]) < 5e-6
    assert len(res[64]["cores"]) > 1000
    assert abs(len(res[32]["cores"]) - len(res[64]["cores"])) <= 0.01 * len(res[64]["cores"])
    assert max_d is not None and max_d <= h


# ---- C3 -------------------------------------------------------------------------------
def test_c3_fp32_energy_tracks_fp64_256():
    """600 renormalised imaginary-time steps of the rotating ground-state problem at 256^3:
    the FP32 run's energy within 1e-6 (relative) of the FP64 run's, psi within 1e-5."""
    n, steps = 256, 600
    pr, spec = initial.ic_spec("c3", n)
    announce(f"C3 {n}^3 rotating imaginary time, {steps} steps, FP64 and FP32")
    res = {}
    for prec in (64, 32):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.init_state(spec)
            r = c.imag_solve(steps, 1e-30, 1.0, 200)
            psi = c.download().astype(np.complex128)
        assert r["steps"] == steps and not r["converged"]
        res[prec] = (r, psi)
        print(f"fp{prec}: E {r['energy']:.10f} mu {r['mu']:.8f} residual "
              f"{r['residual']:.3e}", flush=True)
    e64, e32 = res[64][0]["energy"], res[32][0]["energy"]
    err_e = abs(e32 - e64) / abs(e64)
    err_psi = rel_l2(res[32][1], res[64][1])
    print(f"energy rel err {err_e:.3e}, psi rel L2 {err_psi:.3e}", flush=True)
    assert err_e < 1e-6
    assert err_psi < 1e-5


def test_c3_twin_32_vs_oracle():
    """The same physics and start at 32^3 (C3's box, trap, g, Omega, dtau, phase noise):
    60 FP64 steps against the oracle's renormalised imaginary-time RK3 to 1e-12, and the
    FP32 energy within 1e-6 of the oracle's."""
    pr, psi0 = initial.config_c3(n=32)
    announce("C3 twin 32^3 vs oracle, 60 steps")
    kp = pr.kernel_params()
    s = oracle.Stepper(oracle.params(pr.shape, pr.box, kp["harm_w"], omega_z=kp["omega_r"],
                                     c0=kp["c0"], gnl=kp["gnl"], dt=kp["dt"]), psi0)
    s.imag_rk3(60, 1.0)
    e = s.energy()
    for prec in (64, 32):
        with tbec.Condensate.from_problem(pr, prec=prec) as c:
            c.upload(psi0)
            c.step(60, tbec.IMAG_RK3)
            got = c.download()
            d = c.diagnose()
        err = relmax(got, s.interior)
        err_e = abs(d["energy"] - e["energy"]) / abs(e["energy"])
        print(f"fp{prec}: psi rel max {err:.3e}, energy rel {err_e:.3e}", flush=True)
        if prec == 64:
            assert err < 1e-12 and err_e < 1e-12
        else:
            assert err < 2e-6 and err_e < 1e-6


# ---- C4 -------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4_problems():
    return initial.config_c4()


def test_c4_batch_equals_single_solves_48(c4_problems):
    """All 48 instances (64^3 .. 160^3 and 96x96x192, Omega in [0, 0.8]) solved as one
    FP32 batch for up to 300 steps (stop rule 1e-10), against each instance solved alone
    with the same tile and segment: the same step counts and converged flags, and psi
    and energy bit for bit."""
    probs = c4_problems
    tile, seg, steps = 2, 32, 300
    announce(f"C4 batch of {len(probs)} FP32, {steps} steps, tile {tile} seg {seg}")
    with tbec.Batch([p for p, _ in probs], prec=32) as b:
        b.set_tuning(tile, seg)
        for i, (_, psi0) in enumerate(probs):
            b.upload(i, psi0)
        res = b.imag_solve(steps, 1e-10, 1.0, 50)
        got = [b.download(i) for i in range(len(probs))]
    mismatched = []
    for i, (pr, psi0) in enumerate(probs):
        with tbec.Condensate.from_problem(pr, prec=32) as c:
            c.set_tuning(tile, seg)
            c.upload(psi0)
            r = c.imag_solve(steps, 1e-10, 1.0, 50)
            ref = c.download()
        ok = (r["steps"] == res[i]["steps"] and r["converged"] == res[i]["converged"]
              and r["energy"] == res[i]["energy"] and np.array_equal(got[i], ref))
        if not ok:
            mismatched.append((i, pr.shape, r["steps"], res[i]["steps"]))
    print(f"steps: {[int(r['steps']) for r in res]}", flush=True)
    assert not mismatched, mismatched


def test_c4_heterogeneous_instances_vs_oracle(c4_problems):
    """Three C4 instances of different grid, trap, g and Omega (64^3 rotating at 0.34 with
    g = 3966; 96^3 at Omega = 0.018; 96x96x192 at Omega = 0.31 with g = 395) batched in
    FP64 for 20 steps against the FP64 oracle, each to 1e-12."""
    picks = [0, 8, 14]
    sub = [c4_problems[i] for i in picks]
    announce(f"C4 instances {picks} {[p.shape for p, _ in sub]} FP64 vs oracle, 20 steps")
    with tbec.Batch([p for p, _ in sub], prec=64) as b:
        for i, (_, psi0) in enumerate(sub):
            b.upload(i, psi0)
        b.imag_solve(20, 1e-30, 1.0, 20)
        got = [b.download(i) for i in range(len(sub))]
    for k, (pr, psi0) in enumerate(sub):
        kp = pr.kernel_params()
        s = oracle.Stepper(oracle.params(pr.shape, pr.box, kp["harm_w"],
                                         omega_z=kp["omega_r"], c0=kp["c0"], gnl=kp["gnl"],
                                         dt=kp["dt"]), psi0)
        s.imag_rk3(20, 1.0)
        err = relmax(got[k], s.interior)
        print(f"instance {picks[k]} {pr.shape}: psi rel max {err:.3e}", flush=True)
        assert err < 1e-12


# ---- C5 -------------------------------------------------------------------------------
def _rand_params(dims, seed):
    rng = np.random.default_rng(seed)
    out = []
    for a, b in zip(dims[:-1], dims[1:]):
        out.append(rng.uniform(-1, 1, a * b) * np.sqrt(6.0 / a))
        out.append(rng.uniform(-0.1, 0.1, b))
    return np.concatenate(out).astype(np.float32)


def test_c5_network_forward_and_gradients_vs_oracle():
    """The warm-start network at its production size, 5 -> 4 x 512 (GELU) -> 98,304
    (32^3 x 3 channels), batch 32: forward, loss and every layer's gradient from the
    BF16-MFMA kernels against the BF16-operand, FP64-accumulation oracle (relative L2
    2e-3 forward, 2e-2 gradients, 5e-2 per layer block)."""
    dims = [5, 512, 512, 512, 512, 3 * 32 ** 3]
    announce(f"C5 network {dims} vs oracle")
    rng = np.random.default_rng(13)
    B = 32
    X = rng.uniform(-1, 1, (B, 5)).astype(np.float32)
    T = (0.3 * rng.standard_normal((B, dims[-1]))).astype(np.float32)
    p = _rand_params(dims, 14)
    with tbec.MLP(5, 512, 4, dims[-1], seed=5) as m:
        m.set_params(p)
        y = m.forward(X)
        loss, g = m.loss_grad(X, T)
    loss_o, g_o, y_o = mo.loss_grad(p, dims, X, T)
    print(f"forward rel {rel_l2(y, y_o):.2e} loss {loss:.6e} vs {loss_o:.6e} grad rel "
          f"{rel_l2(g, g_o):.2e}", flush=True)
    assert rel_l2(y, y_o) < 2e-3
    assert loss == pytest.approx(loss_o, rel=5e-3)
    assert rel_l2(g, g_o) < 2e-2
    off = 0
    for a, b in zip(dims[:-1], dims[1:]):
        for nn in (a * b, b):
            assert rel_l2(g[off:off + nn], g_o[off:off + nn]) < 5e-2, off
            off += nn


def test_c5_upsample_32_to_384():
    """tbec_upload_coarse at C5's shapes: a 32^3 x 3 coarse field trilinearly upsampled to
    384^3 on [-12,12)^3 (FP32), against scipy's linear interpolation at 200,000 sampled
    points (a few FP32 ulps, amplified by a/m in the phase division)."""
    from scipy.ndimage import map_coordinates
    cn, n = 32, 384
    announce(f"C5 upsample {cn}^3 -> {n}^3")
    rng = np.random.default_rng(21)
    coarse = np.stack([rng.uniform(0.1, 1.0, (cn,) * 3), rng.uniform(-1, 1, (cn,) * 3),
                       rng.uniform(-1, 1, (cn,) * 3)]).astype(np.float32)
    with tbec.Condensate((n, n, n), (24.0, 24.0, 24.0), (1.0, 1.0, 1.0), prec=32) as c:
        tbec.upload_coarse(c, coarse)
        got = c.download()
    idx = rng.integers(0, n, (3, 200000))
    pos = (idx + 0.5) * cn / n - 0.5
    ch = [map_coordinates(coarse[q].astype(np.float64), pos, order=1, mode="nearest")
          for q in range(3)]
    m = np.hypot(ch[1], ch[2])
    a = np.maximum(ch[0], 0)
    want = a * (ch[1] + 1j * ch[2]) / m
    have = got[idx[0], idx[1], idx[2]]
    err = np.abs(have - want) / (1.0 + a / m)
    print(f"max scaled error {err.max():.3e}", flush=True)
    assert np.all(err < 2e-6)
    assert math.isfinite(float(np.abs(got).max()))
