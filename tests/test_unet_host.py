"""CPU test of the U-Net orchestration (forward, backward, Adam sequencing in
parallelstencil/unet.py) on a float64 host stand-in of the ps_nn_* kernels, against PyTorch
autograd: structural mistakes (a transposed operand, a skip gradient dropped, a wrong layer
order) show up as large errors, rounding does not."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from parallelstencil import unet as U  # noqa: E402
import nn_mock  # noqa: E402


def _torch_ref(net, x, t, scale):
    F = torch.nn.functional
    params = []
    for layer in net.layers():
        k, cin, cout = layer.k, layer.cin, layer.cout
        wt = layer.w.download().reshape(k, k, cin, cout).transpose(3, 2, 0, 1).copy()
        params.append((torch.tensor(wt, requires_grad=True), torch.tensor(layer.b.download().copy(), requires_grad=True)))

    def conv(h, i, relu=True, k=3):
        w, b = params[i]
        h = F.conv2d(h, w, b, padding=k // 2)
        return F.relu(h) if relu else h

    h = torch.tensor(x.transpose(0, 3, 1, 2).copy())
    skips = []
    for l in range(4):
        h = conv(conv(h, 2 * l), 2 * l + 1)
        if l < 3:
            skips.append(h)
            h = F.avg_pool2d(h, 2)
    for l in (2, 1, 0):
        h = torch.cat([F.interpolate(h, scale_factor=2, mode="nearest"), skips[l]], 1)
        h = conv(conv(h, 8 + 2 * l), 9 + 2 * l)
    out = conv(h, 14, relu=False, k=1)
    L = scale * ((out - torch.tensor(t.transpose(0, 3, 1, 2).copy())) ** 2).sum()
    L.backward()
    return out.detach().numpy().transpose(0, 2, 3, 1), float(L), params


def test_unet_gradients_match_autograd(monkeypatch):
    nn_mock.install(U, monkeypatch)
    B, H, W = 2, 16, 24
    net = U.UNet(base=4, seed=5)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((B, H, W, 1))
    t = rng.standard_normal((B, H, W, 3))
    xb = U.Buf(x.shape).upload(x)
    y = net.forward(xb, B, H, W)
    tb, dout, loss = U.Buf(t.shape, "f32").upload(t), U.Buf(t.shape), U.Buf((1,), "f32")
    scale = 1.0 / t.size
    U.lib.ps_nn_mse(y.p, tb.p, None, y.n, scale, dout.p, loss.p, None)
    net.zero_grad()
    net.backward(dout)
    ref, L, params = _torch_ref(net, x, t, scale)
    assert np.allclose(y.download(), ref, rtol=1e-10, atol=1e-12)
    assert abs(loss.download()[0] - L) <= 1e-10 * L
    for layer, (w, b) in zip(net.layers(), params):
        k, cin, cout = layer.k, layer.cin, layer.cout
        gw = layer.gw.download().reshape(k, k, cin, cout).transpose(3, 2, 0, 1)
        assert np.allclose(gw, w.grad.numpy(), rtol=1e-8, atol=1e-12), f"weights of {cin}->{cout}"
        assert np.allclose(layer.gb.download(), b.grad.numpy(), rtol=1e-8, atol=1e-12), f"bias of {cin}->{cout}"
