"""Host (numpy float64) stand-in for the ps_nn_* kernels, for CPU tests of the U-Net
orchestration (parallelstencil/unet.py): buffers live in a dict keyed by fake pointers and
each entry point restates its kernel's definition from include/ps_nn.h."""
import ctypes
import itertools

import numpy as np

_mem = {}
_ids = itertools.count(1 << 20)


def _a(p):
    return _mem[p.value if isinstance(p, ctypes.c_void_p) else p]


class FakeLib:
    def ps_nn_alloc(self, nbytes, out):
        k = next(_ids)
        _mem[k] = np.zeros(max(int(nbytes), 1), np.float64)  # sized generously; views by element count below
        out._obj.value = k
        return 0

    def ps_nn_free(self, p):
        _mem.pop(p.value if isinstance(p, ctypes.c_void_p) else p, None)
        return 0

    def ps_nn_h2d(self, dst, src, nbytes, stream):
        raise NotImplementedError  # Buf.upload / download are patched instead

    def ps_nn_fill_f32(self, x, n, v, stream):
        _a(x)[:n] = v
        return 0

    def ps_nn_gemm(self, M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, bias, epi, splitk, stream):
        a, b, c = _a(A), _a(B), _a(C)
        m, n, k = np.arange(M)[:, None], np.arange(N)[None, :], np.arange(K)
        Am = a[np.arange(M)[:, None] * sam + k[None, :] * sak]
        Bm = b[k[:, None] * sbk + np.arange(N)[None, :] * sbn]
        r = Am @ Bm
        idx = m * ldc + n
        if epi in (1, 2):
            r = r + _a(bias)[:N][None, :]
        if epi == 1:
            r = np.maximum(r, 0)
        if epi == 3:
            c[idx] += r
        else:
            c[idx] = r
        return 0

    def ps_nn_gemm_im2col(self, transpose, M, N, K, x, H, W, C, B, sbk, sbn, Cm, ldc, bias, epi, splitk, stream):
        # the definition in include/ps_nn.h: ps_nn_im2col into a column matrix, then ps_nn_gemm on it
        npix = K if transpose else M
        k = next(_ids)
        _mem[k] = np.zeros(npix * 9 * C)
        self.ps_nn_im2col(x, npix // (H * W), H, W, C, k, stream)
        if transpose:
            r = self.ps_nn_gemm(M, N, K, k, 1, 9 * C, B, sbk, sbn, Cm, ldc, bias, epi, splitk, stream)
        else:
            r = self.ps_nn_gemm(M, N, K, k, 9 * C, 1, B, sbk, sbn, Cm, ldc, bias, epi, splitk, stream)
        del _mem[k]
        return r

    def ps_nn_conv_flip(self, w, Cin, Cout, wf, stream):
        n = 9 * Cin * Cout
        _a(wf)[:n] = _a(w)[:n].reshape(9, Cin, Cout)[::-1].transpose(0, 2, 1).reshape(-1)
        return 0

    def ps_nn_im2col(self, x, B, H, W, C, col, stream):
        X = _a(x)[:B * H * W * C].reshape(B, H, W, C)
        Xp = np.pad(X, ((0, 0), (1, 1), (1, 1), (0, 0)))
        out = np.stack([Xp[:, ky:ky + H, kx:kx + W, :] for ky in range(3) for kx in range(3)], 3)
        _a(col)[:out.size] = out.reshape(-1)
        return 0

    def ps_nn_col2im(self, dcol, B, H, W, C, dx, stream):
        D = _a(dcol)[:B * H * W * 9 * C].reshape(B, H, W, 9, C)
        out = np.zeros((B, H + 2, W + 2, C))
        for kk in range(9):
            ky, kx = kk // 3, kk % 3
            out[:, ky:ky + H, kx:kx + W, :] += D[:, :, :, kk, :]
        _a(dx)[:B * H * W * C] = out[:, 1:H + 1, 1:W + 1, :].reshape(-1)
        return 0

    def ps_nn_pool2(self, x, B, H, W, C, y, stream):
        X = _a(x)[:B * H * W * C].reshape(B, H // 2, 2, W // 2, 2, C)
        _a(y)[:B * H * W * C // 4] = X.mean(axis=(2, 4)).reshape(-1)
        return 0

    def ps_nn_pool2_bwd(self, dy, B, H, W, C, dx, stream):
        D = _a(dy)[:B * (H // 2) * (W // 2) * C].reshape(B, H // 2, 1, W // 2, 1, C)
        _a(dx)[:B * H * W * C] = np.broadcast_to(0.25 * D, (B, H // 2, 2, W // 2, 2, C)).reshape(-1)
        return 0

    def ps_nn_up2(self, x, B, H, W, C, y, stream):
        X = _a(x)[:B * H * W * C].reshape(B, H, 1, W, 1, C)
        _a(y)[:4 * B * H * W * C] = np.broadcast_to(X, (B, H, 2, W, 2, C)).reshape(-1)
        return 0

    def ps_nn_up2_bwd(self, dy, B, H, W, C, dx, stream):
        D = _a(dy)[:4 * B * H * W * C].reshape(B, H, 2, W, 2, C)
        _a(dx)[:B * H * W * C] = D.sum(axis=(2, 4)).reshape(-1)
        return 0

    def ps_nn_concat(self, a, Ca, b, Cb, P, y, stream):
        A, Bv = _a(a)[:P * Ca].reshape(P, Ca), _a(b)[:P * Cb].reshape(P, Cb)
        _a(y)[:P * (Ca + Cb)] = np.concatenate([A, Bv], 1).reshape(-1)
        return 0

    def ps_nn_split(self, dy, Ca, Cb, P, da, db, stream):
        D = _a(dy)[:P * (Ca + Cb)].reshape(P, Ca + Cb)
        _a(da)[:P * Ca] = D[:, :Ca].reshape(-1)
        _a(db)[:P * Cb] = D[:, Ca:].reshape(-1)
        return 0

    def ps_nn_relu_bwd(self, dy, y, n, stream):
        d = _a(dy)
        d[:n] = np.where(_a(y)[:n] > 0, d[:n], 0.0)
        return 0

    def ps_nn_bias_grad(self, dy, M, N, db, stream):
        _a(db)[:N] += _a(dy)[:M * N].reshape(M, N).sum(0)
        return 0

    def ps_nn_mse(self, pred, target, mask, n, scale, dpred, loss, stream):
        d = _a(pred)[:n] - _a(target)[:n]
        mk = _a(mask)[:n] if mask else 1.0
        _a(loss)[0] += scale * float(np.sum(mk * d * d))
        _a(dpred)[:n] = 2 * scale * mk * d
        return 0

    def ps_nn_add_bf16(self, y, x, n, stream):
        _a(y)[:n] += _a(x)[:n]
        return 0

    def ps_nn_adam(self, w, g, m, v, n, lr, b1, b2, eps, c1, c2, w16, stream):
        W_, G, Mm, Vv = _a(w), _a(g), _a(m), _a(v)
        Mm[:n] = b1 * Mm[:n] + (1 - b1) * G[:n]
        Vv[:n] = b2 * Vv[:n] + (1 - b2) * G[:n] ** 2
        W_[:n] = W_[:n] - lr * (Mm[:n] * c1) / (np.sqrt(Vv[:n] * c2) + eps)
        if w16:
            _a(w16)[:n] = W_[:n]
        return 0


def install(U, monkeypatch):
    """Routes parallelstencil.unet through the host stand-in (float64, no BF16 rounding)."""
    fake = FakeLib()
    monkeypatch.setattr(U, "lib", fake)
    monkeypatch.setattr(U, "check", lambda rc: rc)
    monkeypatch.setattr(U.core, "ensure_device", lambda device=0: None)
    monkeypatch.setattr(U.Buf, "upload", lambda self, a: (_a(self.p).__setitem__(slice(0, self.n), np.asarray(a, np.float64).reshape(-1)), self)[1])
    monkeypatch.setattr(U.Buf, "download", lambda self: _a(self.p)[:self.n].reshape(self.shape).copy())
    monkeypatch.setattr(U.Buf, "zero", lambda self: (_a(self.p).__setitem__(slice(0, self.n), 0.0), self)[1])
    return fake
