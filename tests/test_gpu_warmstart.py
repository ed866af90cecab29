"""Warm-start plumbing: coarse-field upsampling on the device against scipy's linear
interpolation, and a small end-to-end warm start that must beat the TF start."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from trapbec_amd import tbec

pytestmark = pytest.mark.gpu


def test_upload_coarse_trilinear():
    cn, shape = 8, (24, 20, 16)
    print(f"\n[case] upload_coarse {cn} -> {shape}", flush=True)
    rng = np.random.default_rng(4)
    coarse = np.stack([rng.uniform(0.1, 1.0, (cn,) * 3), rng.uniform(-1, 1, (cn,) * 3),
                       rng.uniform(-1, 1, (cn,) * 3)]).astype(np.float32)
    with tbec.Condensate(shape, (6.0, 5.0, 4.0), (1.0, 1.0, 1.0), prec=64) as c:
        tbec.upload_coarse(c, coarse)
        got = c.download()
    axes = [(np.arange(n) + 0.5) * cn / n - 0.5 for n in shape]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    ch = [map_coordinates(coarse[q].astype(np.float64), [X, Y, Z], order=1, mode="nearest")
          for q in range(3)]
    m = np.hypot(ch[1], ch[2])
    a = np.maximum(ch[0], 0)
    want = a * (ch[1] + 1j * ch[2]) / m
    # the device interpolates in FP32 (the network's output precision): an absolute
    # error of a few FP32 ulps in each channel, amplified by a/m in the phase division
    assert np.all(np.abs(got - want) < 2e-6 * (1.0 + a / m))


def test_trained_network_beats_tf_start():
    """The whole learned warm start at small scale: ground states of 256 Sobol parameter
    sets from the batched solver (64^3), the production-size MLP trained on them on the
    GPU, then a held-out rotating problem at 128^3.  The TF start runs K imaginary-time
    steps; the network's start must reach that energy and GP residual within K/4 steps."""
    import math
    from trapbec_amd import warmstart as ws
    print("\n[case] warm start end to end, 256 samples 64^3 -> 128^3", flush=True)
    params, X, Y, stats = ws.generate(256, seed=5, n=64, dtau=5e-3, tol=1e-8,
                                      max_steps=8000, chunk=256, log=lambda m: None)
    scale = np.ones(Y.shape[1], np.float32)
    scale[:ws.COARSE ** 3] = 1.0 / Y[:, :ws.COARSE ** 3].max()
    test = np.array([1.0, 0.8, 3000.0, 0.5, 0.0])
    with tbec.MLP(5, 512, 4, 3 * ws.COARSE ** 3, seed=1) as m:
        m.set_data(X, Y * scale)
        m.train(150, 32, 1e-3, seed=2)
        coarse = (m.forward(ws.features(test[None]))[0] / scale).reshape(
            3, ws.COARSE, ws.COARSE, ws.COARSE)
    coarse[0] /= math.sqrt(math.exp(test[4]))
    pr = ws.fine_problem(test, 128, 2e-3)
    K = 16000
    r_tf = ws.refine(pr, ws.tf_start(pr, 1.0, seed=0), 0.0, K)
    r_ml = ws.refine_until(pr, r_tf["final_energy"], r_tf["final_residual"], K, chunk=100,
                           coarse=coarse)
    print(f"TF start after {r_tf['steps']} steps: E {r_tf['final_energy']:.8f} residual "
          f"{r_tf['final_residual']:.3e}; network start after {r_ml['steps']} steps: E "
          f"{r_ml['energy']:.8f} residual {r_ml['residual']:.3e}", flush=True)
    assert r_tf["steps"] == K
    assert r_ml["reached"] and r_ml["steps"] <= K // 4
