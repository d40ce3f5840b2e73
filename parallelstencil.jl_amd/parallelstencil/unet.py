"""U-Net warm start for the pseudo-transient Stokes solver, on the library's own GPU kernels.

No counterpart in the reference (ParallelStencil.jl has no learned component). The network maps
the viscosity field (log10 eta) and the body force to a starting point (Vx, Vy, P) for the PT iteration of
miniapps/Stokes2D.jl; the plain iteration then runs to the same residual tolerance from there,
so the solver alone decides the answer.

Everything that computes runs in libpsamd.so (include/ps_nn.h): BF16 storage, FP32
accumulation, 3x3 convolutions (and their gradients) as MFMA GEMMs that gather the im2col
operand from the activations (no column buffer), average pooling, nearest upsampling,
skip concatenation, ReLU, masked MSE, Adam on FP32 master weights. This module only allocates
buffers, moves host data and sequences the calls.

Architecture: 4 levels (H, H/2, H/4, H/8), `base` channels doubling per level, two 3x3 conv +
ReLU per level on the way down and up, skip concatenation, a 1x1 head with 3 outputs.
Activations are NHWC with h = y and w = x, so a Fortran-ordered field A[ix, iy] is the
row-major plane [iy, ix].
"""
import ctypes
import math

import numpy as np

from . import core
from ._lib import check, lib

EPI_BF16, EPI_BIAS_RELU, EPI_BIAS, EPI_F32_ADD, EPI_F32 = 0, 1, 2, 3, 4


class Buf:
    """Zero-filled device buffer of `shape`, dtype "bf16" or "f32"."""

    def __init__(self, shape, dtype="bf16"):
        core.ensure_device()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.n = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.n * (2 if dtype == "bf16" else 4)
        self.p = ctypes.c_void_p()
        check(lib.ps_nn_alloc(self.nbytes, ctypes.byref(self.p)))

    def upload(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(self.shape))
        host = core.f32_to_bf16(a) if self.dtype == "bf16" else a
        host = np.ascontiguousarray(host)
        check(lib.ps_nn_h2d(self.p, host.ctypes.data_as(ctypes.c_void_p), self.nbytes, None))
        return self

    def download(self):
        host = np.empty(self.shape, dtype=np.uint16 if self.dtype == "bf16" else np.float32)
        check(lib.ps_nn_d2h(host.ctypes.data_as(ctypes.c_void_p), self.p, self.nbytes, None))
        return core.bf16_to_f32(host).reshape(self.shape) if self.dtype == "bf16" else host

    def zero(self):
        if self.dtype == "f32":
            check(lib.ps_nn_fill_f32(self.p, self.n, 0.0, None))
        else:
            check(lib.ps_nn_h2d(self.p, np.zeros(self.n, np.uint16).ctypes.data_as(ctypes.c_void_p), self.nbytes, None))
        return self

    def __del__(self):
        try:
            lib.ps_nn_free(self.p)
        except Exception:
            pass


def gemm(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias=None, epi=EPI_BF16, splitk=1):
    check(lib.ps_nn_gemm(M, N, K, A.p, sam, sak, B.p, sbk, sbn, C.p, ldc, bias.p if bias is not None else None, epi, splitk, None))


def gemm_im2col(transpose, M, N, K, x, H, W, C, B, sbk, sbn, Cm, ldc, bias=None, epi=EPI_BF16, splitk=1):
    """gemm with A the 3x3 im2col of the NHWC input x (never stored): transpose 0 A = col
    (M pixels x K = 9C), transpose 1 A = col^T (the weight gradient's operand)."""
    check(lib.ps_nn_gemm_im2col(transpose, M, N, K, x.p, H, W, C, B.p, sbk, sbn, Cm.p, ldc, bias.p if bias is not None else None, epi,
                                splitk, None))


def _splitk(M, N, K):
    """K chunks for a product whose M x N tiles (64 x 64) alone do not fill the GPU: enough
    chunks for about 1024 workgroups, each chunk at least 256 deep."""
    tiles = -(-M // 64) * -(-N // 64)
    return max(1, min(-(-1024 // tiles), K // 256))


class Conv:
    """3x3 (k=3) or 1x1 (k=1) convolution, stride 1, zero padding; ReLU optional. Weights are
    FP32 masters [k*k*cin, cout] with a BF16 copy used by the GEMMs."""

    def __init__(self, cin, cout, rng, k=3, relu=True):
        self.cin, self.cout, self.k, self.relu = cin, cout, k, relu
        self.K = k * k * cin
        std = math.sqrt(2.0 / self.K)
        w = (rng.standard_normal((self.K, cout)) * std).astype(np.float32)
        self.w = Buf((self.K, cout), "f32").upload(w)
        self.w16 = Buf((self.K, cout), "bf16").upload(w)
        self.b = Buf((cout,), "f32")
        self.gw, self.gb = Buf((self.K, cout), "f32"), Buf((cout,), "f32")
        self.mw, self.vw = Buf((self.K, cout), "f32"), Buf((self.K, cout), "f32")
        self.mb, self.vb = Buf((cout,), "f32"), Buf((cout,), "f32")
        self._bufs = {}

    def params(self):
        return [(self.w, self.gw, self.mw, self.vw, self.w16), (self.b, self.gb, self.mb, self.vb, None)]

    def _buf(self, name, shape):
        b = self._bufs.get(name)
        if b is None or b.shape != tuple(shape):
            b = self._bufs[name] = Buf(shape)
        return b

    def forward(self, x, B, H, W):
        """y = act(conv(x) + b); 3x3 convolutions read their im2col columns straight from x."""
        M = B * H * W
        self.x, self.geom = x, (B, H, W)
        y = self._buf("y", (B, H, W, self.cout))
        epi = EPI_BIAS_RELU if self.relu else EPI_BIAS
        if self.k == 3:
            gemm_im2col(0, M, self.cout, self.K, x, H, W, self.cin, self.w16, self.cout, 1, y, self.cout, self.b, epi)
        else:
            gemm(M, self.cout, self.K, x, self.K, 1, self.w16, self.cout, 1, y, self.cout, self.b, epi)
        self.y = y
        return y

    def backward(self, dy, need_dx=True):
        """dy (bf16, B x H x W x cout) is overwritten by the pre-activation gradient; the weight
        and bias gradients accumulate into gw, gb; returns dx (or None)."""
        B, H, W = self.geom
        M = B * H * W
        if self.relu:
            check(lib.ps_nn_relu_bwd(dy.p, self.y.p, dy.n, None))
        check(lib.ps_nn_bias_grad(dy.p, M, self.cout, self.gb.p, None))
        # gw[kk, co] += sum_pixels col[p, kk] dy[p, co]
        sk = _splitk(self.K, self.cout, M)
        if self.k == 3:
            gemm_im2col(1, self.K, self.cout, M, self.x, H, W, self.cin, dy, self.cout, 1, self.gw, self.cout, epi=EPI_F32_ADD, splitk=sk)
        else:
            gemm(self.K, self.cout, M, self.x, 1, self.K, dy, self.cout, 1, self.gw, self.cout, epi=EPI_F32_ADD, splitk=sk)
        if not need_dx:
            return None
        dx = self._buf("dx", (B, H, W, self.cin))
        if self.k == 1:  # dx[p, ci] = sum_co dy[p, co] w[ci, co]
            gemm(M, self.cin, self.cout, dy, self.cout, 1, self.w16, 1, self.cout, dx, self.cin)
            return dx
        # dx[q, ci] = sum over taps t, co of dy[q - off(t), co] w[t*cin + ci, co]: the 3x3 im2col of
        # dy (off(8-t) = -off(t)) times the weights with reversed taps, each tap transposed
        wf = self._buf("wf", (9 * self.cout, self.cin))
        check(lib.ps_nn_conv_flip(self.w16.p, self.cin, self.cout, wf.p, None))
        gemm_im2col(0, M, self.cin, 9 * self.cout, dy, H, W, self.cout, wf, self.cin, 1, dx, self.cin)
        return dx


class UNet:
    """4-level U-Net (see the module docstring). Input (B, H, W, cin), output (B, H, W, 3); H and W
    divisible by 8."""

    def __init__(self, base=32, cin=1, cout=3, seed=0):
        rng = np.random.default_rng(seed)
        c = [base, 2 * base, 4 * base, 8 * base]
        self.c = c
        self.enc = [(Conv(cin if l == 0 else c[l - 1], c[l], rng), Conv(c[l], c[l], rng)) for l in range(4)]
        self.dec = [(Conv(c[l + 1] + c[l], c[l], rng), Conv(c[l], c[l], rng)) for l in range(3)]
        self.head = Conv(c[0], cout, rng, k=1, relu=False)
        self.cout = cout
        self.t = 0
        self._bufs = {}

    def layers(self):
        out = []
        for a, b in self.enc + self.dec:
            out += [a, b]
        return out + [self.head]

    def _buf(self, name, shape):
        b = self._bufs.get(name)
        if b is None or b.shape != tuple(shape):
            b = self._bufs[name] = Buf(shape)
        return b

    def forward(self, x, B, H, W):
        c = self.c
        hs = [(H >> l, W >> l) for l in range(4)]
        self.geom, self.skips = (B, H, W), []
        h = x
        for l in range(4):
            Hl, Wl = hs[l]
            h = self.enc[l][1].forward(self.enc[l][0].forward(h, B, Hl, Wl), B, Hl, Wl)
            if l < 3:
                self.skips.append(h)
                p = self._buf(f"pool{l}", (B, hs[l + 1][0], hs[l + 1][1], c[l]))
                check(lib.ps_nn_pool2(h.p, B, Hl, Wl, c[l], p.p, None))
                h = p
        for l in (2, 1, 0):
            Hl, Wl = hs[l]
            u = self._buf(f"up{l}", (B, Hl, Wl, c[l + 1]))
            check(lib.ps_nn_up2(h.p, B, hs[l + 1][0], hs[l + 1][1], c[l + 1], u.p, None))
            cat = self._buf(f"cat{l}", (B, Hl, Wl, c[l + 1] + c[l]))
            check(lib.ps_nn_concat(u.p, c[l + 1], self.skips[l].p, c[l], B * Hl * Wl, cat.p, None))
            h = self.dec[l][1].forward(self.dec[l][0].forward(cat, B, Hl, Wl), B, Hl, Wl)
        return self.head.forward(h, B, H, W)

    def backward(self, dout):
        """Gradients of all parameters for the output gradient dout (bf16, B x H x W x cout)."""
        c = self.c
        B, H, W = self.geom
        hs = [(H >> l, W >> l) for l in range(4)]
        dh = self.head.backward(dout)
        dskip = {}
        for l in (0, 1, 2):
            Hl, Wl = hs[l]
            dh = self.dec[l][0].backward(self.dec[l][1].backward(dh))
            du = self._buf(f"dup{l}", (B, Hl, Wl, c[l + 1]))
            ds = self._buf(f"dskip{l}", (B, Hl, Wl, c[l]))
            check(lib.ps_nn_split(dh.p, c[l + 1], c[l], B * Hl * Wl, du.p, ds.p, None))
            dskip[l] = ds
            dl = self._buf(f"dlow{l}", (B, hs[l + 1][0], hs[l + 1][1], c[l + 1]))
            check(lib.ps_nn_up2_bwd(du.p, B, hs[l + 1][0], hs[l + 1][1], c[l + 1], dl.p, None))
            dh = dl
        for l in (3, 2, 1, 0):
            Hl, Wl = hs[l]
            if l < 3:  # dh is the gradient of the pooled output of this level
                dp = self._buf(f"dpool{l}", (B, Hl, Wl, c[l]))
                check(lib.ps_nn_pool2_bwd(dh.p, B, Hl, Wl, c[l], dp.p, None))
                check(lib.ps_nn_add_bf16(dp.p, dskip[l].p, dp.n, None))
                dh = dp
            dh = self.enc[l][0].backward(self.enc[l][1].backward(dh), need_dx=(l > 0))

    def zero_grad(self):
        for layer in self.layers():
            layer.gw.zero()
            layer.gb.zero()

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        c1, c2 = 1.0 / (1.0 - b1 ** self.t), 1.0 / (1.0 - b2 ** self.t)
        for layer in self.layers():
            for w, g, m, v, w16 in layer.params():
                check(lib.ps_nn_adam(w.p, g.p, m.p, v.p, w.n, lr, b1, b2, eps, c1, c2, w16.p if w16 is not None else None, None))

    def weights(self):
        """Host copies of the FP32 master weights, layer by layer: [(w [K, cout], b [cout])]."""
        return [(l.w.download(), l.b.download()) for l in self.layers()]


class WarmStart:
    """Training data handling and inference for the Stokes warm start on an nx x ny grid:
    inputs log10(Mus) and the body force Rog (nx x ny, padded to a multiple of 8 by edge
    replication; the body force marks inclusions whose viscosity equals the matrix's), targets
    Vx, Vy, Pt on the padded grid (padding masked out of the loss), normalised per channel."""

    def __init__(self, nx, ny, base=32, seed=0):
        self.nx, self.ny = nx, ny
        self.H, self.W = _up8(ny + 1), _up8(nx + 1)
        self.net = UNet(base=base, cin=2, seed=seed)
        self.mu = np.zeros(3, np.float32)
        self.sd = np.ones(3, np.float32)
        self.meta = {}

    def _plane(self, A):
        """Fortran field A[ix, iy] -> padded [H, W] plane (edge replication)."""
        P = np.asarray(A, dtype=np.float32).T
        return np.pad(P, ((0, self.H - P.shape[0]), (0, self.W - P.shape[1])), mode="edge")

    def inputs(self, Mus, Rog):
        return np.stack([self._plane(np.log10(Mus)), self._plane(Rog)], -1)

    def targets(self, Vx, Vy, Pt):
        """Pressure is defined up to a constant (the PT iteration drifts it per sample): the
        target is the mean-free pressure."""
        Pt = np.asarray(Pt) - float(np.mean(Pt))
        t = np.stack([self._plane(Vx), self._plane(Vy), self._plane(Pt)], -1)
        m = np.zeros_like(t)
        m[:Vx.shape[1], :Vx.shape[0], 0] = 1
        m[:Vy.shape[1], :Vy.shape[0], 1] = 1
        m[:Pt.shape[1], :Pt.shape[0], 2] = 1
        return t, m

    def save(self, path, meta=None):
        """Writes the grid, the FP32 master weights, the target normalisation and the integers of
        `meta` (a dict, e.g. how it was trained) to `path` (.npz)."""
        arrs = {"grid": np.array([self.nx, self.ny, self.net.c[0]], np.int64), "mu": self.mu, "sd": self.sd}
        for k, v in (meta or {}).items():
            arrs["meta_" + k] = np.array(int(v), np.int64)
        for i, (w, b) in enumerate(self.net.weights()):
            arrs[f"w{i}"], arrs[f"b{i}"] = w, b
        np.savez(path, **arrs)

    @classmethod
    def load(cls, path):
        """A warm start with the weights and normalisation that `save` wrote."""
        with np.load(path) as z:  # plain arrays only (allow_pickle stays off)
            nx, ny, base = (int(v) for v in z["grid"])
            ws = cls(nx, ny, base=base)
            for i, layer in enumerate(ws.net.layers()):
                w, b = z[f"w{i}"], z[f"b{i}"]
                if w.shape != layer.w.shape or b.shape != layer.b.shape:
                    raise ValueError(f"{path}: layer {i} has shape {w.shape}, the network expects {layer.w.shape}")
                layer.w.upload(w)
                layer.w16.upload(w)
                layer.b.upload(b)
            ws.mu[:] = z["mu"]
            ws.sd[:] = z["sd"]
            ws.meta = {k[5:]: int(z[k]) for k in z.files if k.startswith("meta_")}
        return ws

    def set_normalisation(self, T, Mk):
        for c in range(3):
            v = T[..., c][Mk[..., c] > 0]
            self.mu[c], self.sd[c] = float(v.mean()), float(v.std() + 1e-30)

    def train(self, X, T, Mk, epochs, batch=8, lr=1e-3, seed=0, log=None, cosine=False):
        """Adam on the masked MSE of the normalised targets (learning rate constant, or decaying
        as a half cosine over the epochs), in batches of `batch` samples (at most the sample
        count); returns the loss history (per epoch)."""
        lr0 = lr
        n = X.shape[0]
        batch = max(1, min(batch, n))
        H, W = self.H, self.W
        Tn = ((T - self.mu) / self.sd).astype(np.float32)
        xb = Buf((batch, H, W, X.shape[-1]))
        tb, mb = Buf((batch, H, W, 3), "f32"), Buf((batch, H, W, 3), "f32")
        dout, loss = Buf((batch, H, W, 3)), Buf((1,), "f32")
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            if cosine:
                lr = lr0 * 0.5 * (1.0 + math.cos(math.pi * ep / epochs))
            order = rng.permutation(n)
            tot, cnt = 0.0, 0
            for s in range(0, n - batch + 1, batch):
                idx = order[s:s + batch]
                xb.upload(X[idx])
                tb.upload(Tn[idx])
                mb.upload(Mk[idx])
                y = self.net.forward(xb, batch, H, W)
                loss.zero()
                scale = 1.0 / float(Mk[idx].sum())
                check(lib.ps_nn_mse(y.p, tb.p, mb.p, y.n, scale, dout.p, loss.p, None))
                self.net.zero_grad()
                self.net.backward(dout)
                self.net.adam(lr)
                tot += float(loss.download()[0])
                cnt += 1
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        return hist

    def predict(self, Mus, Rog):
        """(Vx, Vy, Pt) host arrays on the solver's staggered shapes."""
        H, W = self.H, self.W
        xb = self._in if getattr(self, "_in", None) is not None else Buf((1, H, W, 2))
        self._in = xb
        xb.upload(self.inputs(Mus, Rog)[None])
        y = self.net.forward(xb, 1, H, W).download()[0]
        y = y * self.sd + self.mu
        nx, ny = self.nx, self.ny
        Vx = np.asfortranarray(y[:ny, :nx + 1, 0].T, dtype=np.float64)
        Vy = np.asfortranarray(y[:ny + 1, :nx, 1].T, dtype=np.float64)
        Pt = np.asfortranarray(y[:ny, :nx, 2].T, dtype=np.float64)
        return Vx, Vy, Pt


def _up8(n):
    return ((n + 7) // 8) * 8


def generate_samples(solver, seeds, log10_contrast=1.0):
    """Cold PT solves (to the solver's tolerance) of the inclusion samples `seeds`; returns
    (list of (Mus, Rog, Vx, Vy, Pt), list of (iterations, seconds, residual))."""
    import time
    from .workloads import stokes2d_inclusions
    data, stats = [], []
    for s in seeds:
        Mus, Rog = stokes2d_inclusions(solver.nx, solver.ny, s, log10_contrast=log10_contrast)
        solver.set_material(Mus, Rog)
        solver.reset()
        core.synchronize()
        t0 = time.perf_counter()
        it, err = solver.solve()
        t1 = time.perf_counter()
        stats.append((it, t1 - t0, err))
        data.append((Mus, Rog, solver.Vx.to_host(), solver.Vy.to_host(), solver.Pt.to_host()))
    return data, stats


def generate_samples_batch(bsolver, seeds, log10_contrast=1.0, log=None, every=64):
    """generate_samples on a Stokes2DBatch: the solves run S at a time through its slot queue,
    with the same results per seed (iterations, residual and fields bit for bit those of the
    single-sample solver). Seconds per seed are the batch's wall time shared in proportion to
    the seed's iterations. `log` receives a line every `every` finished solves."""
    import time
    from .workloads import stokes2d_inclusions
    seeds = list(seeds)
    res = [None] * len(seeds)
    core.synchronize()
    t0 = time.perf_counter()
    mats = (stokes2d_inclusions(bsolver.nx, bsolver.ny, s, log10_contrast=log10_contrast) for s in seeds)
    kept = {}

    def samples():
        for k, m in enumerate(mats):
            kept[k] = m
            yield m

    for k, (i, it, err, Vx, Vy, Pt) in enumerate(bsolver.solve_queue(samples())):
        Mus, Rog = kept.pop(i)
        res[i] = (it, err, (Mus, Rog, Vx, Vy, Pt))
        if log and (k + 1) % every == 0:
            log(f"stokes {bsolver.nx}^2 batch of {bsolver.S}: {k + 1}/{len(seeds)} solves, {time.perf_counter() - t0:.1f} s")
    wall = time.perf_counter() - t0
    total = max(sum(r[0] for r in res), 1)
    return [r[2] for r in res], [(r[0], wall * r[0] / total, r[1]) for r in res]


def mirror_x(sample):
    """A sample (Mus, Rog, Vx, Vy, Pt) mirrored in x: the miniapp's problem is symmetric about
    x = lx/2 (free-slip walls, the body force along y), so the mirrored fields solve the mirrored
    inclusion setup, with Vx (on the x faces) changing sign."""
    Mus, Rog, Vx, Vy, Pt = sample
    f = lambda A: np.asfortranarray(np.asarray(A)[::-1, :])  # noqa: E731
    return f(Mus), f(Rog), np.asfortranarray(-np.asarray(Vx)[::-1, :]), f(Vy), f(Pt)


def dataset(ws, data):
    X = np.stack([ws.inputs(d[0], d[1]) for d in data]).astype(np.float32)
    TM = [ws.targets(d[2], d[3], d[4]) for d in data]
    T = np.stack([t for t, _ in TM]).astype(np.float32)
    Mk = np.stack([m for _, m in TM]).astype(np.float32)
    return X, T, Mk


def resampled_predictor(ws, n, n2, lx=10.0):
    """Warm start on an n2 x n2 grid from the network trained on n x n: log10 of the viscosity and
    the body force are interpolated (bilinear) onto the n-grid's nodes, the network predicts there, and the predicted
    Vx, Vy, Pt are interpolated back onto the n2-grid's staggered positions (miniapps/Stokes2D.jl:
    nodes at (i-1)dx, Vx between them at (i-3/2)dx, Vy at (j-3/2)dy, dx = lx/(n-1))."""
    from scipy.ndimage import map_coordinates

    def resample(A, src_pos, dst_pos):
        # A sampled at positions src_pos[d][i] (uniform per axis) -> values at dst_pos grid
        coords = []
        for d in range(2):
            s0, ds = src_pos[d][0], src_pos[d][1] - src_pos[d][0]
            coords.append((dst_pos[d] - s0) / ds)
        X, Y = np.meshgrid(coords[0], coords[1], indexing="ij")
        return np.asfortranarray(map_coordinates(np.asarray(A, dtype=np.float64), [X, Y], order=1, mode="nearest"))

    def axes(m):
        dx = lx / (m - 1)
        nodes = np.arange(m) * dx
        faces = (np.arange(m + 1) - 0.5) * dx
        return nodes, faces

    n_nodes, n_faces = axes(n)
    m_nodes, m_faces = axes(n2)

    def predict(Mus2, Rog2):
        logm = resample(np.log10(Mus2), (m_nodes, m_nodes), (n_nodes, n_nodes))
        rog = resample(Rog2, (m_nodes, m_nodes), (n_nodes, n_nodes))
        Vx, Vy, Pt = ws.predict(10.0 ** logm, rog)
        return (resample(Vx, (n_faces, n_nodes), (m_faces, m_nodes)), resample(Vy, (n_nodes, n_faces), (m_nodes, m_faces)),
                resample(Pt, (n_nodes, n_nodes), (m_nodes, m_nodes)))
    return predict


def _pred_err(pred, sol):
    """Relative L2 error of predicted (Vx, Vy, Pt) against a solution (pressures mean-free)."""
    out = {}
    for name, p, r in zip(("Vx", "Vy", "P"), pred, sol):
        p, r = np.asarray(p, np.float64), np.asarray(r, np.float64)
        if name == "P":
            p, r = p - p.mean(), r - r.mean()
        den = float(np.sqrt(np.sum(r * r)))
        out[name] = float(np.sqrt(np.sum((p - r) ** 2)) / den) if den > 0 and np.all(np.isfinite(p)) else float("inf")
    return out


def evaluate(ws, solver, seeds, log10_contrast=1.0, predict=None, log=None):
    """Per held-out seed: (cold iterations, cold seconds, warm iterations, warm seconds including
    inference and the host/device copies of the prediction, final residuals, relative L2 error of
    the prediction against the converged cold solution per field, pressure mean-free); with a
    solver that records its residual history, the cold and warm histories (one residual per
    check). `predict` maps (viscosity, body force) to (Vx, Vy, Pt) (default: the network at the
    solver's resolution). A seed whose plain iteration does not converge has no reference to
    compare with: it is recorded with its cold run only (err_cold above the tolerance). `log`
    receives one line per seed."""
    predict = predict or ws.predict
    import time
    from .workloads import stokes2d_inclusions
    out = []
    seeds = list(seeds)
    for k, s in enumerate(seeds):
        Mus, Rog = stokes2d_inclusions(solver.nx, solver.ny, s, log10_contrast=log10_contrast)
        solver.set_material(Mus, Rog)
        solver.reset()
        core.synchronize()
        t0 = time.perf_counter()
        itc, errc = solver.solve()
        t1 = time.perf_counter()
        hist_cold = list(getattr(solver, "history", None) or [])
        if not errc <= solver.eps:
            out.append(dict(seed=s, cold_iters=itc, cold_s=t1 - t0, err_cold=errc))
            if log:
                log(f"held-out {solver.nx}^2 {k + 1}/{len(seeds)} (seed {s}): plain iteration not converged after {itc} iterations")
            continue
        sol = (solver.Vx.to_host(), solver.Vy.to_host(), solver.Pt.to_host())
        solver.reset()
        core.synchronize()
        t2 = time.perf_counter()
        Vx, Vy, Pt = predict(Mus, Rog)
        fallback = not (np.all(np.isfinite(Vx)) and np.all(np.isfinite(Vy)) and np.all(np.isfinite(Pt)))
        itw, errw, hist_warm = 0, float("inf"), []
        if not fallback:  # the warm-started attempt gets at most twice the plain iteration count
            solver.reset(Vx, Vy, Pt)
            itmax = solver.itermax
            solver.itermax = max(2 * itc, 2 * solver.nout)
            try:
                itw, errw = solver.solve()
            finally:
                solver.itermax = itmax
            hist_warm = list(getattr(solver, "history", None) or [])
            fallback = not errw <= solver.eps
        if fallback:  # the plain iteration from rest decides; its cost is charged to the warm start
            solver.reset()
            it2, errw = solver.solve()
            itw += it2
        t3 = time.perf_counter()
        rec = dict(seed=s, cold_iters=itc, cold_s=t1 - t0, warm_iters=itw, warm_s=t3 - t2, err_cold=errc, err_warm=errw,
                   fallback=fallback, pred_err=_pred_err((Vx, Vy, Pt), sol))
        if hist_cold:
            rec["hist_cold"] = [float(f"{e:.4e}") for e in hist_cold]
            rec["hist_warm"] = [float(f"{e:.4e}") for e in hist_warm]
        out.append(rec)
        if log:
            log(f"held-out {solver.nx}^2 {k + 1}/{len(seeds)} (seed {s}): cold {itc} iterations / {t1 - t0:.3f} s, "
                f"warm {itw} iterations / {t3 - t2:.3f} s")
    return out
