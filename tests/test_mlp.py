"""The surrogate network: the oracle's back-propagation against finite differences
(CPU), BF16 rounding known answers (CPU), and the GPU's BF16 MFMA forward and
backward kernels against the oracle with the same rounding (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def small_net(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in mo.SIZES:
        fan_in = shape[1] if len(shape) == 2 else mo.HID
        d[name] = rng.uniform(-1, 1, shape) * np.sqrt(3.0 / fan_in)
    return mo.flatten(d)


def test_bf16_rounding_known_answers():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even gives 1
    assert mo.bf16(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even gives 1 + 2^-6
    assert mo.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert mo.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert mo.bf16(-3.140625) == -3.140625
    assert mo.bf16(65280.0) == 65280.0


def test_oracle_gradient_matches_finite_differences():
    p = small_net(1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(24, mo.NIN))
    t = rng.normal(size=24)
    _, g = mo.loss_and_grad(p, x, t, bf=False)
    idx = np.concatenate([rng.choice(mo.N_PARAMS, 60, replace=False),
                          [mo.N_PARAMS - 1, mo.N_PARAMS - 2, 0, 2048]])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mo.loss_and_grad(pp, x, t)[0] - mo.loss_and_grad(pm, x, t)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle(wg):
    print("case: MLP forward/gradient vs oracle, 200 rows", flush=True)
    from watergap2_amd.surrogate import MLP
    net = MLP(seed=3)
    p = small_net(4).astype(np.float32)
    net.set_params(p)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(200, mo.NIN)).astype(np.float32)
    t = rng.normal(size=200).astype(np.float32)
    y = net.forward(x)
    y_ref, _ = mo.forward(p.astype(np.float64), x.astype(np.float64), bf=True)
    assert rel_err(y, y_ref) < 2e-3, rel_err(y, y_ref)
    loss, g = net.gradient(x, t)
    loss_ref, g_ref = mo.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss - loss_ref) <= 2e-3 * loss_ref
    o = 0
    for name, shape in mo.SIZES:
        n = int(np.prod(shape))
        e = rel_err(g[o:o + n], g_ref[o:o + n])
        assert e < 2e-2, (name, e)
        o += n


@pytest.mark.gpu
def test_gpu_training_fits_a_smooth_function(wg):
    print("case: MLP training on a synthetic regression", flush=True)
    from watergap2_amd.surrogate import MLP
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, size=(8192, mo.NIN)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] ** 2 - 0.5 * x[:, 2] * x[:, 3]).astype(np.float32)
    net = MLP(seed=1)
    losses = net.train(x, t, epochs=30, batch=256, lr=2e-3, seed=3)
    assert losses[-1] < 0.05 * losses[0], losses
    y = net.forward(x[:512])
    assert float(np.mean((y - t[:512]) ** 2)) < 0.05 * float(np.var(t))
