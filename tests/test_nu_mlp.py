"""Learned Nu predictor: the bf16 matrix-core perceptron against a plain
PyTorch reference of the same arithmetic (bf16-rounded weights and layer
inputs, sums in float64), its loss gradient against the same reference's
explicit backward pass, and that Adam training fits a smooth target."""
import numpy as np
import pytest
import torch

import axisem3d_amd as ax
from axisem3d_amd import _abi
from tests.helpers import rel_l2


def _bf(t):
    return t.to(torch.bfloat16).to(torch.float64)


def _ref(mlp, params, x, y=None, under_weight=1.0):
    """Forward (and, with targets, the loss gradient in parameter layout)."""
    layers = [(torch.tensor(W, dtype=torch.float64), torch.tensor(b, dtype=torch.float64))
              for W, b in mlp.layers(params)]
    a = torch.tensor(x, dtype=torch.float64)
    acts = [a]
    for l, (W, b) in enumerate(layers):
        z = _bf(a) @ _bf(W).T + b
        if l + 1 < len(layers):
            a = torch.relu(z)
            acts.append(a)
    z = z[:, 0]
    if y is None:
        return z.numpy()
    yt = torch.tensor(y, dtype=torch.float64)
    d = z - yt
    w = torch.where(d < 0, torch.full_like(d, under_weight), torch.ones_like(d))
    dZ = (2.0 * w * d / len(y))[:, None]
    grads = []
    for l in reversed(range(len(layers))):
        W, _ = layers[l]
        grads.insert(0, (_bf(dZ).T @ _bf(acts[l]), dZ.sum(0)))
        if l > 0:
            dZ = (_bf(dZ) @ _bf(W)) * (acts[l] > 0)
    return np.concatenate([np.concatenate([gw.reshape(-1).numpy(), gb.numpy()]) for gw, gb in grads])


@pytest.mark.gpu
def test_mlp_forward_matches_reference():
    print("case mlp forward", flush=True)
    mlp = ax.NuMLP(seed=5)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((300, _abi.NFEAT)).astype(np.float32)  # 300: not a multiple of the tile
    p = mlp.params()
    y = mlp.predict(x)
    yr = _ref(mlp, p, x)
    err = rel_l2(y, yr)
    print("forward rel L2 %.3e" % err, flush=True)
    assert np.abs(yr).max() > 0.1
    assert err < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("under_weight", [1.0, 4.0])
def test_mlp_gradient_matches_reference(under_weight):
    print("case mlp gradient uw=%g" % under_weight, flush=True)
    mlp = ax.NuMLP(seed=6)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((200, _abi.NFEAT)).astype(np.float32)
    y = rng.standard_normal(200).astype(np.float32)
    p = mlp.params()
    g = mlp.gradient(x, y, under_weight)
    gr = _ref(mlp, p, x, y, under_weight)
    o = 0
    for din, dout in zip(mlp.dims[:-1], mlp.dims[1:]):
        for n in (din * dout, dout):
            e = rel_l2(g[o:o + n], gr[o:o + n])
            print("layer block %d..%d rel L2 %.3e" % (o, o + n, e), flush=True)
            assert e < 1e-4
            o += n
    assert o == mlp.nparams


@pytest.mark.gpu
def test_mlp_training_fits_smooth_target():
    print("case mlp training", flush=True)
    rng = np.random.default_rng(3)

    def target(x):
        return np.sin(2.0 * x[:, 0]) + 0.5 * x[:, 1] ** 2 - x[:, 2] * x[:, 3]

    x = rng.uniform(-1, 1, (8192, _abi.NFEAT)).astype(np.float32)
    xt = rng.uniform(-1, 1, (2048, _abi.NFEAT)).astype(np.float32)
    y, yt = target(x).astype(np.float32), target(xt).astype(np.float32)
    mlp = ax.NuMLP(seed=7)
    mse0 = float(np.mean((mlp.predict(xt) - yt) ** 2))
    mlp.set_data(x, y)
    mlp.train(epochs=40, batch=256, lr=1e-3, seed=1)
    mse = float(np.mean((mlp.predict(xt) - yt) ** 2))
    print("held-out mse %.4f -> %.4f (target variance %.4f)" % (mse0, mse, float(np.var(yt))), flush=True)
    assert mse < 0.05 * float(np.var(yt))
