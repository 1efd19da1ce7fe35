"""The guess network's hand-written kernels against a float64 CPU reference
of the same network (PyTorch autograd on the host).

The reference rounds what the kernels round -- the BF16 weights, inputs and
hidden activations -- so the remaining differences are FP32 accumulation
order on the matrix cores and the BF16 rounding of the backward operands
(the SiLU gradients).  Bar: forward outputs within 1e-3 of the output scale,
every gradient block within 2e-2 of its largest magnitude (BF16 has an
8-bit mantissa, 2^-8 = 3.9e-3 relative).
"""
import numpy as np
import pytest

from engine_path import ngmi

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu


def bf16(a):
    t = torch.as_tensor(np.asarray(a, dtype=np.float32))
    return t.to(torch.bfloat16).to(torch.float64)


def reference(net, p, X, T):
    W1, b1, W2, b2, W3, b3, w4, b4 = [torch.as_tensor(np.array(x), dtype=torch.float64)
                                      for x in net.unpack(p)]
    W1b, W2b, W3b, w4b = (bf16(W1.numpy()), bf16(W2.numpy()), bf16(W3.numpy()), bf16(w4.numpy()))
    for t in (W1b, b1, W2b, b2, W3b, b3, w4b, b4):
        t.requires_grad_(True)
    Xp = torch.zeros(X.shape[0], 32, dtype=torch.float64)
    Xp[:, :16] = bf16(X)
    silu = torch.nn.functional.silu
    z1 = Xp @ W1b.T + b1
    h1 = silu(z1).to(torch.bfloat16).to(torch.float64)
    z2 = h1 @ W2b.T + b2
    h2 = silu(z2).to(torch.bfloat16).to(torch.float64)
    z3 = h2 @ W3b.T + b3
    h3 = silu(z3).to(torch.bfloat16).to(torch.float64)
    y = h3 @ w4b + b4
    loss = ((y - torch.as_tensor(T, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    grads = [W1b.grad, b1.grad, W2b.grad, b2.grad, W3b.grad, b3.grad, w4b.grad, b4.grad]
    return y.detach().numpy(), float(loss), [g.numpy().reshape(-1) for g in grads]


@pytest.mark.parametrize('M', [64, 1000])
def test_forward_and_gradients(M):
    print('case: guess network M=%d' % M, flush=True)
    rng = np.random.default_rng(M)
    net = ngmi.GuessNet(seed=7)
    X = rng.normal(size=(M, 16)).astype(np.float32)
    T = rng.normal(scale=0.1, size=M).astype(np.float32)
    dX = torch.from_numpy(X).cuda()
    dT = torch.from_numpy(T).cuda()
    dy = torch.zeros(M, dtype=torch.float32, device='cuda')
    net.forward(dX.data_ptr(), dy.data_ptr(), M)
    torch.cuda.synchronize()
    p = net.params()
    y_ref, loss_ref, g_ref = reference(net, p, X, T)
    y = dy.cpu().numpy()
    ferr = np.abs(y - y_ref).max() / max(np.abs(y_ref).max(), 1e-30)
    print('forward scaled err %.2e' % ferr, flush=True)
    assert ferr < 1e-3
    loss = net.train_step(dX.data_ptr(), dT.data_ptr(), M, 0.0)
    assert abs(loss - loss_ref) <= 1e-4 * max(loss_ref, 1e-30) + 1e-7
    g = net.unpack(net.params(grads=True))
    names = ['W1', 'b1', 'W2', 'b2', 'W3', 'b3', 'w4', 'b4']
    for name, gg, gr in zip(names, g, g_ref):
        gg = np.asarray(gg, dtype=np.float64).reshape(-1)
        scale = max(np.abs(gr).max(), 1e-30)
        err = np.abs(gg - gr).max() / scale
        print('grad %-3s scaled err %.2e (scale %.2e)' % (name, err, scale), flush=True)
        assert err < 2e-2, name


def test_training_reduces_loss():
    """Adam on a smooth synthetic target: the loss falls by 10x."""
    rng = np.random.default_rng(3)
    M = 4096
    X = rng.normal(size=(M, 16)).astype(np.float32)
    T = (0.3 * np.tanh(X[:, 0] - 0.5 * X[:, 1]) + 0.1 * X[:, 2] * X[:, 3]).astype(np.float32)
    net = ngmi.GuessNet(seed=11)
    dX = torch.from_numpy(X).cuda()
    dT = torch.from_numpy(T).cuda()
    first = net.train_step(dX.data_ptr(), dT.data_ptr(), M, 1e-3)
    for _ in range(300):
        last = net.train_step(dX.data_ptr(), dT.data_ptr(), M, 1e-3, want_loss=False)
    last = net.train_step(dX.data_ptr(), dT.data_ptr(), M, 0.0)
    print('loss %.3e -> %.3e' % (first, last), flush=True)
    assert last < 0.1 * first
