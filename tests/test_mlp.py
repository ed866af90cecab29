"""The warm-start MLP: the numpy oracle pinned by finite differences (CPU), and libtbec's
BF16-MFMA forward/backward against that oracle (GPU)."""
import numpy as np
import pytest

import mlp_oracle as mo


def _rand_params(dims, seed):
    rng = np.random.default_rng(seed)
    out = []
    for a, b in zip(dims[:-1], dims[1:]):
        out.append(rng.uniform(-1, 1, a * b) * np.sqrt(6.0 / a))
        out.append(rng.uniform(-0.1, 0.1, b))
    return np.concatenate(out).astype(np.float32)


def test_gelu_derivative_finite_difference():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.abs(fd - mo.gelu_grad(z)).max() < 1e-8
    assert mo.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mo.gelu(np.array([10.0]))[0] - 10.0) < 1e-12


def test_oracle_backward_finite_difference():
    """With BF16 rounding off the oracle's gradient equals central differences of its
    own FP64 loss: pins the backward pass (transposes, GELU', bias sums, 2/n)."""
    dims = [5, 7, 6, 4]
    p = _rand_params(dims, 1).astype(np.float64)
    rng = np.random.default_rng(2)
    X, T = rng.standard_normal((3, 5)), rng.standard_normal((3, 4))
    _, g, _ = mo.loss_grad(p, dims, X, T, round_bf16=False)
    h = 1e-6
    for i in range(0, p.size, 7):
        q = p.copy()
        q[i] += h
        lp = mo.loss_grad(q, dims, X, T, False)[0]
        q[i] -= 2 * h
        lm = mo.loss_grad(q, dims, X, T, False)[0]
        assert abs((lp - lm) / (2 * h) - g[i]) < 1e-7 * max(1.0, abs(g[i])), i


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3])
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 1e-3) < 1e-3 * 2 ** -8


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [[5, 64, 64, 96], [5, 512, 512, 512, 512, 3 * 16 ** 3]])
def test_gpu_forward_and_gradients_vs_oracle(dims):
    """Forward output and every parameter gradient from the GPU kernels agree with the
    BF16-operand FP64-accumulation oracle (FP32 accumulation on the device)."""
    from trapbec_amd import tbec
    print(f"\n[case] mlp dims {dims}", flush=True)
    rng = np.random.default_rng(3)
    B = 64
    X = rng.uniform(-1, 1, (B, dims[0])).astype(np.float32)
    T = rng.standard_normal((B, dims[-1])).astype(np.float32) * 0.3
    p = _rand_params(dims, 4)
    with tbec.MLP(dims[0], dims[1], len(dims) - 2, dims[-1], seed=5) as m:
        m.set_params(p)
        y = m.forward(X)
        loss, g = m.loss_grad(X, T)
    loss_o, g_o, y_o = mo.loss_grad(p, dims, X, T)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    print(f"forward rel {rel(y, y_o):.2e} loss {loss:.6e} vs {loss_o:.6e} "
          f"grad rel {rel(g, g_o):.2e}")
    assert rel(y, y_o) < 2e-3
    assert loss == pytest.approx(loss_o, rel=5e-3)
    assert rel(g, g_o) < 2e-2
    # per-layer check so a transposed or misplaced block cannot hide in the total
    off = 0
    for a, b in zip(dims[:-1], dims[1:]):
        for n in (a * b, b):
            assert rel(g[off:off + n], g_o[off:off + n]) < 5e-2, off
            off += n


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    """A few hundred Adam steps fit a smooth synthetic map far below the initial loss."""
    from trapbec_amd import tbec
    print("\n[case] mlp training", flush=True)
    rng = np.random.default_rng(6)
    X = rng.uniform(-1, 1, (512, 5)).astype(np.float32)
    W = rng.standard_normal((5, 96)).astype(np.float32)
    T = np.sin(X @ W).astype(np.float32)
    with tbec.MLP(5, 128, 3, 96, seed=7) as m:
        m.set_data(X, T)
        hist = m.train(60, 64, 2e-3, seed=8)
    print(f"loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    assert hist[-1] < 0.1 * hist[0]
