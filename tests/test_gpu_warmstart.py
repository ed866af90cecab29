"""Warm-start plumbing: coarse-field upsampling on the device against scipy's linear
interpolation, and a small end-to-end warm start that must beat the TF start."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates

from trapbec_amd import initial, tbec

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
