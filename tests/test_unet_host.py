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


def test_warm_start_save_load_round_trip(monkeypatch, tmp_path):
    nn_mock.install(U, monkeypatch)
    ws = U.WarmStart(15, 15, base=4, seed=2)
    ws.mu[:] = [0.1, -0.2, 0.3]
    ws.sd[:] = [1.5, 2.5, 3.5]
    f = tmp_path / "ws.npz"
    ws.save(f)
    wl = U.WarmStart.load(f)
    assert (wl.nx, wl.ny, wl.net.c[0]) == (15, 15, 4)
    assert np.array_equal(wl.mu, ws.mu) and np.array_equal(wl.sd, ws.sd)
    for (w0, b0), (w1, b1), layer in zip(ws.net.weights(), wl.net.weights(), wl.net.layers()):
        assert np.array_equal(w0, w1) and np.array_equal(b0, b1)
        assert np.allclose(layer.w16.download(), w1, rtol=2 ** -8)  # the BF16 copy the GEMMs use
    with pytest.raises(ValueError):  # a file of another network
        np.savez(f, grid=np.array([15, 15, 4]), mu=ws.mu, sd=ws.sd, **{f"w{i}": np.zeros((1, 1), np.float32) for i in range(15)},
                 **{f"b{i}": np.zeros(1, np.float32) for i in range(15)})
        U.WarmStart.load(f)


def test_mirrored_stokes_samples_solve_the_mirrored_problem():
    """The x-mirrored training samples: 300 PT iterations of the oracle (miniapps/Stokes2D.jl) on an
    off-centre inclusion setup and on its mirror give mirrored fields, Vx with its sign changed
    (to rounding: the 2-D averages sum their four neighbours in mirrored order). A mirror along
    y, a missing sign or a shifted staggered index fails by orders of magnitude."""
    from oracle import ps_oracle as O
    from parallelstencil import workloads as W
    nx = ny = 31
    lx = 10.0
    dx = dy = lx / (nx - 1)

    def run(Mus, Rog, iters=300, Vdmp=4.0):
        f = lambda *sh: np.asfortranarray(np.zeros(sh))  # noqa: E731
        a = dict(Pt=f(nx, ny), dtPt=f(nx, ny), divV=f(nx, ny), Vx=f(nx + 1, ny), Vy=f(nx, ny + 1), txx=f(nx, ny), tyy=f(nx, ny),
                 txy=f(nx - 1, ny - 1), Rx=f(nx - 1, ny - 2), Ry=f(nx - 2, ny - 1), dVxdt=f(nx - 1, ny - 2), dVydt=f(nx - 2, ny - 1),
                 dtVx=f(nx - 1, ny - 2), dtVy=f(nx - 2, ny - 1))
        O.stokes2d_timesteps(a["dtVx"], a["dtVy"], a["dtPt"], Mus, 1.0, 0.25, min(dx, dy) ** 2, max(nx, ny))
        for _ in range(iters):
            O.stokes2d_p(a["divV"], a["Pt"], a["Vx"], a["Vy"], a["dtPt"], dx, dy)
            O.stokes2d_tau(a["divV"], a["txx"], a["tyy"], a["txy"], a["Vx"], a["Vy"], Mus, dx, dy)
            O.stokes2d_dv(a["Rx"], a["Ry"], a["dVxdt"], a["dVydt"], a["Pt"], Rog, a["txx"], a["tyy"], a["txy"], 1 - Vdmp / nx,
                          1 - Vdmp / ny, dx, dy)
            O.stokes2d_v(a["Vx"], a["Vy"], a["dVxdt"], a["dVydt"], a["dtVx"], a["dtVy"])
            O.bc_y(a["Vx"])
            O.bc_x(a["Vy"])
        return a["Vx"], a["Vy"], a["Pt"]

    Mus, Rog = W.stokes2d_inclusions(nx, ny, 7, log10_contrast=1.0)
    Mus, Rog = np.asfortranarray(Mus, dtype=np.float64), np.asfortranarray(Rog, dtype=np.float64)
    assert not np.array_equal(Mus, Mus[::-1, :])  # the setup is not symmetric itself
    Vx, Vy, Pt = run(Mus, Rog)
    mMus, mRog, mVx, mVy, mPt = U.mirror_x((Mus, Rog, Vx, Vy, Pt))
    Vx2, Vy2, Pt2 = run(np.asfortranarray(mMus), np.asfortranarray(mRog))
    for got, want in ((Vx2, mVx), (Vy2, mVy), (Pt2, mPt)):
        scale = np.max(np.abs(want))
        assert scale > 0 and np.max(np.abs(got - want)) <= 1e-12 * scale


def test_network_entry_points_refuse_negative_sizes():
    # a pair of negative sizes multiplies to a positive element count: refused before any launch
    import pytest
    from parallelstencil import _lib
    L = _lib.lib
    calls = [
        ("ps_nn_conv_flip", (None, -2, -3, None, None)),
        ("ps_nn_im2col", (None, 1, -4, -4, 2, None, None)),
        ("ps_nn_col2im", (None, -1, 4, 4, -2, None, None)),
        ("ps_nn_pool2", (None, 1, -4, 4, 2, None, None)),
        ("ps_nn_pool2_bwd", (None, -1, 4, 4, -2, None, None)),
        ("ps_nn_up2", (None, 1, -4, -4, 2, None, None)),
        ("ps_nn_up2_bwd", (None, 1, 4, -4, -2, None, None)),
        ("ps_nn_concat", (None, -3, None, 5, 10, None, None)),
        ("ps_nn_split", (None, 5, -3, 10, None, None, None)),
    ]
    for name, args in calls:
        with pytest.raises(_lib.ArgumentError, match=f"{name}: negative size"):
            _lib.check(getattr(L, name)(*args))
