"""ORACLE -- test infrastructure, not product code.

FP64 NumPy restatement of the warm-start network of
binder_amd/csrc/warmstart.hip (forward, MSE loss and the full backward pass),
used only by tests/ to check the HIP kernels.  The network has no
counterpart in the reference (BINDER always starts its pyramid from the
identity, BINDER/FastRegistration.py:206-225); what pins this restatement:

* conv3d: scipy.ndimage.correlate of each (output, input) channel pair;
* pooling / up-sampling backward: the adjoint (dot-product) identity;
* the head: scipy.fft.dctn(norm="ortho") restricted to the lowest C x C x C
  (zero at frequencies past the grid);
* the whole backward pass: central finite differences of the FP64 loss.

``bf16=True`` rounds, like the HIP path, every stored activation and
gradient tensor and the convolution weights to bfloat16 (round to nearest
even), so the two can be compared closely; accumulation stays FP64.

Parameter vector layout (identical to the HIP handle): per conv layer
W[cout][27][cin] (tap t = 9(dx+1) + 3(dy+1) + (dz+1)) then bias[cout]; then
the head W[3][64], bias[3].
"""
import numpy as np

CONV_LAYERS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64), (64, 64),
               (64, 64), (64, 64), (112, 64), (64, 64)]
A1, B1, A2, B2, A3, B3, A4, B4, C4, D4, A5, B5 = range(12)
NCOEF = 8  # default coefficient count per axis (a network's own is ncoef)


def round_bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split_params(p):
    """[(W [cout,27,cin], b [cout]) per conv layer], (Wh [3,64], bh [3])."""
    p = np.asarray(p, np.float64)
    out, off = [], 0
    for ci, co in CONV_LAYERS:
        W = p[off:off + co * 27 * ci].reshape(co, 27, ci)
        off += co * 27 * ci
        b = p[off:off + co]
        off += co
        out.append((W, b))
    Wh = p[off:off + 192].reshape(3, 64)
    bh = p[off + 192:off + 195]
    return out, (Wh, bh)


def num_params():
    return sum(co * 27 * ci + co for ci, co in CONV_LAYERS) + 195


def _taps():
    for t in range(27):
        yield t, t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1


def conv3d(x, W, b):
    """y[b,x,y,z,co] = sum_{t,ci} x[b, x+dx, y+dy, z+dz, ci] W[co,t,ci] + b[co]
    with zero padding; x [B,X,Y,Z,Cin]."""
    B, X, Y, Z, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, X, Y, Z, W.shape[0]))
    for t, dx, dy, dz in _taps():
        y += xp[:, 1 + dx:1 + dx + X, 1 + dy:1 + dy + Y, 1 + dz:1 + dz + Z, :] @ W[:, t, :].T
    return y + b


def conv3d_backward(x, W, g):
    """Gradients of conv3d for the output gradient g: (dx, dW, db)."""
    B, X, Y, Z, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    g2 = g.reshape(-1, g.shape[-1])
    for t, dx, dy, dz in _taps():
        sl = (slice(None), slice(1 + dx, 1 + dx + X), slice(1 + dy, 1 + dy + Y),
              slice(1 + dz, 1 + dz + Z), slice(None))
        dW[:, t, :] = g2.T @ xp[sl].reshape(-1, x.shape[-1])
        dxp[sl] += g @ W[:, t, :]
    return dxp[:, 1:-1, 1:-1, 1:-1, :], dW, g2.sum(0)


def pool(x):
    B, X, Y, Z, C = x.shape
    return x.reshape(B, X // 2, 2, Y // 2, 2, Z // 2, 2, C).mean(axis=(2, 4, 6))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_backward(g):
    B, X, Y, Z, C = g.shape
    return g.reshape(B, X // 2, 2, Y // 2, 2, Z // 2, 2, C).sum(axis=(2, 4, 6))


def dct_rows(n, k=NCOEF):
    """Rows a < k of the orthonormal DCT-II matrix of length n; rows at or
    past n (frequencies the grid does not have) are zero."""
    C = np.zeros((k, n))
    a = np.arange(min(k, n))[:, None]
    i = np.arange(n)[None, :]
    C[:min(k, n)] = np.sqrt(2.0 / n) * np.cos(np.pi * (2 * i + 1) * a / (2 * n))
    C[0] = np.sqrt(1.0 / n)
    return C


def head_project(U, ncoef=NCOEF):
    """U [B,3,hx,hy,hz] -> coef [B,3,C,C,C] = DCT-II_ortho(U)[:C,:C,:C] / sqrt(h)."""
    hx, hy, hz = U.shape[2:]
    Cx, Cy, Cz = (dct_rows(h, ncoef) for h in (hx, hy, hz))
    c = np.einsum("ai,bj,ck,snijk->snabc", Cx, Cy, Cz, U, optimize=True)
    return c / np.sqrt(hx * hy * hz)


def head_expand(dc, shape):
    hx, hy, hz = shape
    ncoef = dc.shape[-1]
    Cx, Cy, Cz = (dct_rows(h, ncoef) for h in (hx, hy, hz))
    dU = np.einsum("ai,bj,ck,snabc->snijk", Cx, Cy, Cz, dc, optimize=True)
    return dU / np.sqrt(hx * hy * hz)


def make_input(fixed_bins, moving_bins, L=32, K=32, bf16=True):
    """[B,X,Y,Z,8]: channels bin/(L-1), bin/(K-1), then zeros (the division
    in float32, as the HIP path)."""
    f = np.asarray(fixed_bins, np.float32) * np.float32(1.0 / (L - 1))
    m = np.asarray(moving_bins, np.float32) * np.float32(1.0 / (K - 1))
    x = np.zeros(f.shape + (8,))
    x[..., 0] = f
    x[..., 1] = m
    return round_bf16(x) if bf16 else x


class Net:
    def __init__(self, params, bf16=True, ncoef=NCOEF):
        self.bf16 = bf16
        self.ncoef = ncoef
        self.conv, (self.Wh, self.bh) = split_params(params)
        if bf16:
            self.conv = [(round_bf16(W), b) for W, b in self.conv]
        self.q = round_bf16 if bf16 else (lambda a: a)

    def forward(self, x0):
        q, cv = self.q, self.conv
        c = {"x0": x0}

        def layer(i, x):
            return q(np.maximum(conv3d(x, *cv[i]), 0.0))

        c["t1a"] = layer(A1, x0)
        c["s1"] = layer(B1, c["t1a"])
        c["p1"] = q(pool(c["s1"]))
        c["t2a"] = layer(A2, c["p1"])
        c["s2"] = layer(B2, c["t2a"])
        c["p2"] = q(pool(c["s2"]))
        c["t3a"] = layer(A3, c["p2"])
        c["s3"] = layer(B3, c["t3a"])
        c["p3"] = q(pool(c["s3"]))
        c["t4a"] = layer(A4, c["p3"])
        c["t4b"] = layer(B4, c["t4a"])
        c["t4c"] = layer(C4, c["t4b"])
        c["s4"] = layer(D4, c["t4c"])
        c["cat"] = np.concatenate([up(c["s4"]), c["s3"]], axis=-1)
        c["t5a"] = layer(A5, c["cat"])
        c["s5"] = layer(B5, c["t5a"])
        U = np.einsum("bxyzc,nc->bnxyz", c["s5"], self.Wh) + self.bh[None, :, None, None, None]
        c["U"] = U
        coef = head_project(U, self.ncoef)
        self.cache = c
        return coef

    def backward(self, dcoef):
        """Parameter gradient (flat, the handle's layout) for dL/dcoef."""
        q, cv, c = self.q, self.conv, self.cache
        grads = [None] * len(cv)
        dU = head_expand(dcoef, c["U"].shape[2:])
        dWh = np.einsum("bnxyz,bxyzc->nc", dU, c["s5"])
        dbh = dU.sum(axis=(0, 2, 3, 4))
        g = q(np.einsum("bnxyz,nc->bxyzc", dU, self.Wh) * (c["s5"] > 0))

        def back(i, x, g, mask):
            dx, dW, db = conv3d_backward(x, cv[i][0], g)
            grads[i] = (dW, db)
            if mask is None:
                return q(dx)
            return q(dx * (mask > 0))

        g = back(B5, c["t5a"], g, c["t5a"])
        dcat = back(A5, c["cat"], g, None)
        g = q(up_backward(dcat[..., :64]) * (c["s4"] > 0))
        g = back(D4, c["t4c"], g, c["t4c"])
        g = back(C4, c["t4b"], g, c["t4b"])
        g = back(B4, c["t4a"], g, c["t4a"])
        dp3 = back(A4, c["p3"], g, None)
        g = q((pool_backward(dp3) + dcat[..., 64:]) * (c["s3"] > 0))
        g = back(B3, c["t3a"], g, c["t3a"])
        dp2 = back(A3, c["p2"], g, None)
        g = q(pool_backward(dp2) * (c["s2"] > 0))
        g = back(B2, c["t2a"], g, c["t2a"])
        dp1 = back(A2, c["p1"], g, None)
        g = q(pool_backward(dp1) * (c["s1"] > 0))
        g = back(B1, c["t1a"], g, c["t1a"])
        back(A1, c["x0"], g, None)
        flat = []
        for dW, db in grads:
            flat += [dW.ravel(), db]
        flat += [dWh.ravel(), dbh]
        return np.concatenate(flat)


def mse(coef, targets):
    """Mean squared error over all coefficients and its gradient."""
    e = np.asarray(coef, np.float64) - np.asarray(targets, np.float64)
    return float(np.mean(e * e)), 2.0 * e / e.size
