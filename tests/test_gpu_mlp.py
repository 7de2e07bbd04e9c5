"""The overlap MLP of the learned lambda schedule (q_amd/csrc/qmi_mlp.hip):
forward pass and gradients against a PyTorch FP32 CPU reference of the same
network (weights rounded to BF16 as the kernels use them), and training
convergence."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def torch_net(mlp, flat):
    W1, b1, W2, b2, W3, b3 = [torch.tensor(a, dtype=torch.float32) for a in mlp.split(flat)]
    # the kernels use BF16 weights (FP32 biases)
    W1, W2, W3 = [w.to(torch.bfloat16).to(torch.float32) for w in (W1, W2, W3)]
    return [t.clone().requires_grad_(True) for t in (W1, b1, W2, b2, W3, b3)]


def torch_forward(params, x):
    W1, b1, W2, b2, W3, b3 = params
    h1 = torch.nn.functional.silu(x @ W1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
    return torch.sigmoid(h2 @ W3.T + b3).reshape(-1)


@pytest.fixture(scope="module")
def net():
    from q_amd.mlp import MLP
    m = MLP(35, 128, seed=7)
    yield m
    m.close()


def test_forward_matches_torch(net):
    rng = np.random.default_rng(0)
    x = rng.normal(size=(300, 35)).astype(np.float32)
    y = net.forward(x)
    params = torch_net(net, net.params)
    ref = torch_forward(params, torch.tensor(x)).detach().numpy()
    print("forward max abs err", np.abs(y - ref).max())
    assert np.abs(y - ref).max() < 2e-5


def test_gradient_matches_torch_autograd(net):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(257, 35)).astype(np.float32)
    t = rng.uniform(size=257).astype(np.float32)
    loss, g = net.grad(x, t)
    params = torch_net(net, net.params)
    out = torch_forward(params, torch.tensor(x))
    ref_loss = ((out - torch.tensor(t)) ** 2).mean()
    ref_loss.backward()
    ref = np.concatenate([p.grad.numpy().reshape(-1) for p in params])
    print("loss", loss, float(ref_loss), "grad max abs err", np.abs(g - ref).max(), "max |grad|", np.abs(ref).max())
    assert abs(loss - float(ref_loss)) < 1e-6 * max(1.0, abs(float(ref_loss)))
    assert np.abs(g - ref).max() < 1e-4 * np.abs(ref).max()
    # every block of the parameter vector is exercised
    for a, b in zip(net.split(g), net.split(ref.astype(np.float32))):
        assert np.allclose(a, b, rtol=1e-3, atol=1e-6 * np.abs(ref).max())


def test_training_reduces_loss():
    from q_amd.mlp import MLP
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4096, 8)).astype(np.float32)
    t = (1 / (1 + np.exp(-(x[:, 0] * 2 - x[:, 1] + 0.5 * x[:, 2] * x[:, 3])))).astype(np.float32)
    m = MLP(8, 128, seed=3)
    l0, _ = m.grad(x, t)
    l1 = m.train(x, t, epochs=40, batch=256, lr=2e-3, seed=4)
    l2, _ = m.grad(x, t)
    print("loss before", l0, "last epoch", l1, "after", l2)
    assert l2 < 0.1 * l0
    # reproducible: the same seed gives the same weights
    m2 = MLP(8, 128, seed=3)
    m2.train(x, t, epochs=40, batch=256, lr=2e-3, seed=4)
    assert np.array_equal(m.params, m2.params)
    m.close()
    m2.close()
