"""Kernels of the Stokes warm-start U-Net (BF16 MFMA GEMM, im2col/col2im, pooling, upsampling,
concatenation, masked MSE, Adam) on the GPU against CPU references: numpy float64 GEMMs of the
same BF16 inputs, and a PyTorch (CPU, float64) re-implementation of the network with the same
weights for the forward output and every parameter gradient."""
import numpy as np
import pytest

import parallelstencil as ps
from parallelstencil import unet as U

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ps.init_parallel_stencil(ps.AMDGPU, ps.Float64, 2, module=__name__)


def _bf(a):
    return ps.bf16_to_f32(ps.f32_to_bf16(np.asarray(a, np.float32))).astype(np.float64)


@pytest.mark.parametrize("mode", ["nn", "tn", "nt"])
@pytest.mark.parametrize("splitk", [1, 3])
def test_gemm_against_numpy(mode, splitk):
    rng = np.random.default_rng(1)
    M, N, K = 100, 70, 200
    A = _bf(rng.standard_normal((M, K)))
    B = _bf(rng.standard_normal((K, N)))
    if mode == "nn":  # A row-major [M,K], B row-major [K,N]
        a, b, sam, sak, sbk, sbn = A, B, K, 1, N, 1
    elif mode == "tn":  # A stored [K,M]
        a, b, sam, sak, sbk, sbn = A.T.copy(), B, 1, M, N, 1
    else:  # B stored [N,K]
        a, b, sam, sak, sbk, sbn = A, B.T.copy(), K, 1, 1, K
    da, db = U.Buf(a.shape).upload(a), U.Buf(b.shape).upload(b)
    C = U.Buf((M, N), "f32")
    U.gemm(M, N, K, da, sam, sak, db, sbk, sbn, C, N, epi=U.EPI_F32_ADD if splitk > 1 else U.EPI_F32, splitk=splitk)
    ref = A @ B
    assert np.abs(C.download() - ref).max() <= 1e-5 * np.abs(ref).max()
    # bf16 epilogue with bias + relu
    bias = rng.standard_normal(N).astype(np.float32)
    bb = U.Buf((N,), "f32").upload(bias)
    Y = U.Buf((M, N))
    U.gemm(M, N, K, da, sam, sak, db, sbk, sbn, Y, N, bias=bb, epi=U.EPI_BIAS_RELU)
    r = np.maximum(ref + bias, 0)
    assert np.abs(Y.download() - r).max() <= 8e-3 * np.abs(r).max()


def test_data_movement_kernels_against_numpy():
    """im2col / col2im, pooling, upsampling, concat / split against their definitions."""
    import nn_mock
    rng = np.random.default_rng(2)
    B, H, W, C = 2, 6, 10, 5
    x = _bf(rng.standard_normal((B, H, W, C)))
    xb = U.Buf(x.shape).upload(x)
    ref = nn_mock.FakeLib()
    def host(name, *args, out_n):
        """Runs the host definition of kernel `name` on numpy inputs; returns its output."""
        nn_mock._mem.clear()
        keys = []
        for a in args:
            if isinstance(a, np.ndarray):
                k = len(nn_mock._mem) + 1
                nn_mock._mem[k] = a.reshape(-1).astype(np.float64).copy()
                keys.append(k)
            else:
                keys.append(a)
        nn_mock._mem[999] = np.zeros(out_n)
        getattr(ref, name)(*keys, 999, None)
        return nn_mock._mem[999]
    col = U.Buf((B * H * W, 9 * C))
    ps.core.check(U.lib.ps_nn_im2col(xb.p, B, H, W, C, col.p, None))
    assert np.array_equal(col.download().reshape(-1), host("ps_nn_im2col", x, B, H, W, C, out_n=col.n))
    dcol = _bf(rng.standard_normal(col.shape))
    dcb = U.Buf(col.shape).upload(dcol)
    dx = U.Buf(x.shape)
    ps.core.check(U.lib.ps_nn_col2im(dcb.p, B, H, W, C, dx.p, None))
    r = host("ps_nn_col2im", dcol, B, H, W, C, out_n=dx.n)
    assert np.abs(dx.download().reshape(-1) - r).max() <= 8e-3 * np.abs(r).max()
    y = U.Buf((B, H // 2, W // 2, C))
    ps.core.check(U.lib.ps_nn_pool2(xb.p, B, H, W, C, y.p, None))
    r = host("ps_nn_pool2", x, B, H, W, C, out_n=y.n)
    assert np.abs(y.download().reshape(-1) - r).max() <= 8e-3 * np.abs(r).max()
    u = U.Buf((B, 2 * H, 2 * W, C))
    ps.core.check(U.lib.ps_nn_up2(xb.p, B, H, W, C, u.p, None))
    assert np.array_equal(u.download().reshape(-1), host("ps_nn_up2", x, B, H, W, C, out_n=u.n))
    ps.core.check(U.lib.ps_nn_up2_bwd(u.p, B, H, W, C, dx.p, None))
    assert np.array_equal(dx.download(), 4 * x)  # sums of four equal bf16 values are exact
    cat = U.Buf((B, H, W, 2 * C))
    ps.core.check(U.lib.ps_nn_concat(xb.p, C, dx.p, C, B * H * W, cat.p, None))
    c = cat.download().reshape(-1, 2 * C)
    assert np.array_equal(c[:, :C], x.reshape(-1, C)) and np.array_equal(c[:, C:], 4 * x.reshape(-1, C))
    a2, b2 = U.Buf(x.shape), U.Buf(x.shape)
    ps.core.check(U.lib.ps_nn_split(cat.p, C, C, B * H * W, a2.p, b2.p, None))
    assert np.array_equal(a2.download(), x) and np.array_equal(b2.download(), 4 * x)


def _torch_unet(net, x, base):
    """The same network in torch (float64), weights copied from `net`; returns the output and
    the list of parameter tensors (w [cout, cin, k, k], b) in net.layers() order."""
    F = torch.nn.functional
    params = []
    for layer in net.layers():
        w, b = layer.w.download(), layer.b.download()
        k, cin, cout = layer.k, layer.cin, layer.cout
        # [(ky*k+kx)*cin + ci, co] -> [co, ci, ky, kx]; the bf16 copy is what the GPU multiplies with
        wt = _bf(w).reshape(k, k, cin, cout).transpose(3, 2, 0, 1).copy()
        params.append((torch.tensor(wt, requires_grad=True), torch.tensor(b.astype(np.float64), requires_grad=True)))

    def conv(h, i, relu=True, k=3):  # i: index in net.layers(): enc 0..7, dec (level l) 8+2l, 9+2l, head 14
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
    return out, params


def test_unet_forward_and_gradients_against_torch():
    base, B, H, W = 8, 2, 16, 24
    net = U.UNet(base=base, seed=3)
    rng = np.random.default_rng(7)
    x = _bf(rng.standard_normal((B, H, W, 1)))
    xb = U.Buf(x.shape).upload(x)
    y = net.forward(xb, B, H, W)
    t = rng.standard_normal((B, H, W, 3)).astype(np.float32)
    tb = U.Buf(t.shape, "f32").upload(t)
    dout, loss = U.Buf(t.shape), U.Buf((1,), "f32")
    scale = 1.0 / t.size
    ps.core.check(U.lib.ps_nn_mse(y.p, tb.p, None, y.n, scale, dout.p, loss.p, None))
    net.zero_grad()
    net.backward(dout)

    ref, params = _torch_unet(net, x, base)
    ref_np = ref.detach().numpy().transpose(0, 2, 3, 1)
    got = y.download()
    assert np.linalg.norm(got - ref_np) <= 3e-2 * np.linalg.norm(ref_np)
    L = scale * ((ref - torch.tensor(t.transpose(0, 3, 1, 2).astype(np.float64))) ** 2).sum()
    L.backward()
    assert abs(float(loss.download()[0]) - float(L)) <= 3e-2 * float(L)
    for layer, (w, b) in zip(net.layers(), params):
        k, cin, cout = layer.k, layer.cin, layer.cout
        gw = layer.gw.download().reshape(k, k, cin, cout).transpose(3, 2, 0, 1)
        rw = w.grad.numpy()
        # BF16 activations and gradients: the direction agrees closely, the norm within rounding
        # amplified by cancellation in the pixel sums (the float64 host test pins the structure)
        cos = float(np.sum(gw * rw) / (np.linalg.norm(gw) * np.linalg.norm(rw)))
        assert cos >= 0.99 and np.linalg.norm(gw - rw) <= 0.15 * np.linalg.norm(rw), f"weight gradient of a {cin}->{cout} conv"
        gb, rb = layer.gb.download(), b.grad.numpy()
        assert np.linalg.norm(gb - rb) <= 0.15 * np.linalg.norm(rb) + 1e-6


def test_unet_training_reduces_loss():
    """A few hundred Adam steps on a smooth synthetic map (input -> 3 smoothed copies) cut the loss."""
    B, H, W = 4, 32, 32
    net = U.UNet(base=8, seed=1)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((16, H, W, 1)).astype(np.float32)
    k = np.ones(5) / 5
    sm = lambda a: np.apply_along_axis(lambda r: np.convolve(r, k, "same"), 1, np.apply_along_axis(lambda r: np.convolve(r, k, "same"), 0, a))
    T = np.stack([np.stack([sm(x[..., 0]), -sm(x[..., 0]), 0.5 * x[..., 0]], -1) for x in X]).astype(np.float32)
    ws = U.WarmStart.__new__(U.WarmStart)
    ws.net, ws.H, ws.W = net, H, W
    ws.mu, ws.sd = np.zeros(3, np.float32), np.ones(3, np.float32)
    hist = ws.train(X, T, np.ones_like(T), epochs=40, batch=B, lr=1e-2)
    print("loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.2 * hist[0]


@pytest.mark.parametrize("C,misalign", [(2, 0), (5, 0), (32, 0), (64, 0), (32, 1)])
def test_implicit_im2col_gemm_matches_explicit(C, misalign):
    """ps_nn_gemm_im2col reads the 3x3 im2col columns straight from the NHWC input: bit for bit
    ps_nn_im2col + ps_nn_gemm, for the convolution (transpose 0, both epilogues) and its weight
    gradient (transpose 1, split-K); a misaligned input takes the element-by-element loads."""
    import ctypes
    rng = np.random.default_rng(C + 100 * misalign)
    B, H, W, N = 2, 6, 10, 48
    P, K9 = B * H * W, 9 * C
    x = _bf(rng.standard_normal((B, H, W, C)))
    xs = U.Buf((P * C + misalign,))
    xs.upload(np.concatenate([np.zeros(misalign), x.reshape(-1)]))
    xp = ctypes.c_void_p(xs.p.value + 2 * misalign)
    col = U.Buf((P, K9))
    ps.core.check(U.lib.ps_nn_im2col(xp, B, H, W, C, col.p, None))
    w = U.Buf((K9, N)).upload(_bf(rng.standard_normal((K9, N))))
    bias = U.Buf((N,), "f32").upload(rng.standard_normal(N).astype(np.float32))
    for epi, dt in ((U.EPI_F32, "f32"), (U.EPI_BIAS_RELU, "bf16")):
        ye, yi = U.Buf((P, N), dt), U.Buf((P, N), dt)
        b = bias if epi == U.EPI_BIAS_RELU else None
        U.gemm(P, N, K9, col, K9, 1, w, N, 1, ye, N, bias=b, epi=epi)
        ps.core.check(U.lib.ps_nn_gemm_im2col(0, P, N, K9, xp, H, W, C, w.p, N, 1, yi.p, N, b.p if b else None, epi, 1, None))
        assert np.array_equal(ye.download(), yi.download()), (C, epi)
        if epi == U.EPI_F32 and not misalign:  # the convolution itself, in float64
            xpad = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
            cref = np.concatenate([xpad[:, dy:dy + H, dx:dx + W, :] for dy in range(3) for dx in range(3)], axis=3)
            ref = cref.reshape(P, K9) @ w.download().astype(np.float64)
            assert np.abs(yi.download() - ref).max() <= 1e-5 * np.abs(ref).max()
    dy = U.Buf((P, N)).upload(_bf(rng.standard_normal((P, N))))
    for sk in (1, 3):
        ge, gi = U.Buf((K9, N), "f32").zero(), U.Buf((K9, N), "f32").zero()
        U.gemm(K9, N, P, col, 1, K9, dy, N, 1, ge, N, epi=U.EPI_F32_ADD, splitk=sk)
        ps.core.check(U.lib.ps_nn_gemm_im2col(1, K9, N, P, xp, H, W, C, dy.p, N, 1, gi.p, N, None, U.EPI_F32_ADD, sk, None))
        if sk == 1:
            assert np.array_equal(ge.download(), gi.download()), (C, sk)
        else:  # split-K partial sums meet in atomic adds of either order
            g = ge.download()
            assert np.abs(gi.download() - g).max() <= 1e-6 * np.abs(g).max(), (C, sk)
    with pytest.raises(Exception, match="sizes do not match"):
        ps.core.check(U.lib.ps_nn_gemm_im2col(0, P, N, K9 + 1, xp, H, W, C, w.p, N, 1, col.p, N, None, U.EPI_F32, 1, None))


@pytest.mark.parametrize("cin,cout", [(3, 5), (32, 32), (16, 64)])
def test_implicit_data_gradient(cin, cout):
    """dx of a 3x3 convolution as im2col(dy) . wf (ps_nn_conv_flip: taps reversed, each tap
    transposed; an exact permutation) against the adjoint in float64: dcol = dy . W^T summed back
    by col2im; dx is BF16 from FP32 sums."""
    rng = np.random.default_rng(cin * cout)
    B, H, W = 2, 6, 10
    P = B * H * W
    w = _bf(rng.standard_normal((9 * cin, cout)))
    dyh = _bf(rng.standard_normal((P, cout)))
    wb, dyb = U.Buf(w.shape).upload(w), U.Buf(dyh.shape).upload(dyh)
    wf = U.Buf((9 * cout, cin))
    ps.core.check(U.lib.ps_nn_conv_flip(wb.p, cin, cout, wf.p, None))
    assert np.array_equal(wf.download(), w.reshape(9, cin, cout)[::-1].transpose(0, 2, 1).reshape(9 * cout, cin))
    dx = U.Buf((P, cin))
    U.gemm_im2col(0, P, cin, 9 * cout, dyb, H, W, cout, wf, cin, 1, dx, cin)
    dcol = (dyh @ w.T).reshape(B, H, W, 9, cin)
    ref = np.zeros((B, H + 2, W + 2, cin))
    for t in range(9):
        ref[:, t // 3:t // 3 + H, t % 3:t % 3 + W, :] += dcol[:, :, :, t, :]
    ref = ref[:, 1:H + 1, 1:W + 1, :].reshape(P, cin)
    assert np.abs(dx.download() - ref).max() <= 8e-3 * np.abs(ref).max()


def test_implicit_im2col_gemm_at_network_size():
    # a full-resolution layer of the network (1 x 512^2, 32 -> 32 channels): the forward and the
    # weight gradient bit for bit those of the stored column matrix
    rng = np.random.default_rng(7)
    B, H, W, C, N = 1, 512, 512, 32, 32
    P, K9 = B * H * W, 9 * C
    x = U.Buf((P, C)).upload(rng.standard_normal((P, C)).astype(np.float32))
    col = U.Buf((P, K9))
    ps.core.check(U.lib.ps_nn_im2col(x.p, B, H, W, C, col.p, None))
    w = U.Buf((K9, N)).upload(rng.standard_normal((K9, N)).astype(np.float32))
    bias = U.Buf((N,), "f32").upload(rng.standard_normal(N).astype(np.float32))
    ye, yi = U.Buf((P, N)), U.Buf((P, N))
    U.gemm(P, N, K9, col, K9, 1, w, N, 1, ye, N, bias=bias, epi=U.EPI_BIAS_RELU)
    U.gemm_im2col(0, P, N, K9, x, H, W, C, w, N, 1, yi, N, bias, U.EPI_BIAS_RELU)
    assert np.array_equal(ye.download(), yi.download())
    dy = U.Buf((P, N)).upload(rng.standard_normal((P, N)).astype(np.float32))
    ge, gi = U.Buf((K9, N), "f32").zero(), U.Buf((K9, N), "f32").zero()
    U.gemm(K9, N, P, col, 1, K9, dy, N, 1, ge, N, epi=U.EPI_F32_ADD, splitk=1)
    U.gemm_im2col(1, K9, N, P, x, H, W, C, dy, N, 1, gi, N, epi=U.EPI_F32_ADD, splitk=1)
    assert np.array_equal(ge.download(), gi.download())
