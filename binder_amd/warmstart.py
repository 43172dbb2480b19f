"""Learned warm start of the 3D MI registration.

A 3D convolutional encoder-decoder (binder_amd/csrc/warmstart.hip; channels
16/32/48/64, 3x3x3 convolutions, BF16 weights and activations with FP32
accumulation) reads the binned fixed and moving image at the half-resolution
pyramid level and predicts the C x C x C lowest orthonormal DCT-II
coefficients (C = ncoef, at most 16; 16 for the shipped network) of each
component of the full-resolution displacement.  It is trained on the
GPU with its own forward and backward kernels on seeded synthetic pairs,
each labelled by the field the plain (identity-start) pyramid converges to.

The prediction only chooses where the plain EM starts: FastRegistration
runs the same EM to convergence from it (``FastRegistration.run_warm``), and
falls back to the plain cold pyramid whenever the warm start's objective is
worse than the identity start's.  The reference has no counterpart; its
pyramid always starts from the identity (BINDER/FastRegistration.py:206-225).

Coefficient convention (``field_coefficients``): for the displacement
u = F - x of a field F on an Nx x Ny x Nz grid,
coef[n, a, b, c] = sum_ijk C_Nx[a, i] C_Ny[b, j] C_Nz[c, k] u_n[i, j, k] / sqrt(Nx Ny Nz),
with C_N the orthonormal DCT-II matrix, a, b, c < C.
"""
import ctypes
import os

import numpy as np

NCOEF = 16  # coefficients per axis of new networks and labels


class _Api:
    """ctypes entry points of the network, bound on first use so that the
    host-only helpers below (and worker processes that only build inputs)
    never load the HIP library."""
    _fns = None

    @classmethod
    def get(cls):
        if cls._fns is None:
            from . import _core
            vp, ci, P, sig = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER, _core._sig
            cls._fns = dict(
                create=sig("binder_warmstart_create", ci, P(vp), ci, ci, ci, ci, ci,
                           ctypes.c_ulonglong),
                destroy=sig("binder_warmstart_destroy", None, vp),
                nparam=sig("binder_warmstart_num_params", ci, vp, P(ctypes.c_longlong)),
                get=sig("binder_warmstart_get_params", ci, vp, vp),
                set=sig("binder_warmstart_set_params", ci, vp, vp),
                data=sig("binder_warmstart_set_dataset", ci, vp, ci, vp, vp, vp, ci, ci),
                train=sig("binder_warmstart_train", ci, vp, vp, ci, ci, ctypes.c_double, vp),
                grad=sig("binder_warmstart_gradients", ci, vp, vp, ci, P(ctypes.c_double), vp),
                pred_data=sig("binder_warmstart_predict_dataset", ci, vp, vp, ci, vp),
                pred=sig("binder_warmstart_predict", ci, vp, ci, vp, vp, ci, ci, vp),
                check=_core._check, ptr=_core._ptr)
        return cls._fns


# (cin, cout) of the 3x3x3 convolutions, in parameter-vector order
CONV_LAYERS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64), (64, 64),
               (64, 64), (64, 64), (112, 64), (64, 64)]
HEAD_IN = 64


def num_params():
    n = sum(co * 27 * ci + co for ci, co in CONV_LAYERS)
    return n + 3 * HEAD_IN + 3


def network_grid(full_shape):
    """Input grid of the network for a full-resolution shape: the
    half-resolution pyramid level, ceil(N / 2) per axis, padded up to a
    multiple of 8 (at least 32)."""
    out = []
    for n in full_shape:
        half = -(-int(n) // 2)
        out.append(max(32, -(-half // 8) * 8))
    return tuple(out)


def pair_inputs(full_binned_voxels, full_binned_nodes, grid, check=True):
    """Network inputs of one pair: the half-resolution pyramid level's bins
    (full[::2, ::2, ::2], as FastRegistration.move_resolution(2) takes them)
    zero-padded into `grid`, uint8.  check=False skips the [0, 255] range
    test (for callers whose bin counts are at most 256)."""
    out = []
    for b in (full_binned_voxels, full_binned_nodes):
        sub = np.asarray(b)[::2, ::2, ::2]
        if check and (sub.min() < 0 or sub.max() > 255):
            raise ValueError("bins must be in [0, 255]")
        a8 = sub.astype(np.uint8)
        if sub.shape == tuple(grid):
            a = a8
        else:
            a = np.zeros(grid, np.uint8)
            a[:sub.shape[0], :sub.shape[1], :sub.shape[2]] = a8
        out.append(a)
    return out[0], out[1]


def dct_rows(n, k=NCOEF):
    """Rows a < k of the orthonormal DCT-II matrix of length n, [k][n]
    (rows past n are zero)."""
    C = np.zeros((k, n))
    a = np.arange(min(k, n))[:, None]
    i = np.arange(n)[None, :]
    C[:min(k, n)] = np.sqrt(2.0 / n) * np.cos(np.pi * (2 * i + 1) * a / (2 * n))
    C[0] = np.sqrt(1.0 / n)
    return C


def field_coefficients(F, positions=None, ncoef=NCOEF):
    """Normalised low-frequency coefficients [3, C, C, C] (C = ncoef) of the
    displacement F - positions of a field F [Nx, Ny, Nz, 3] (positions
    default to the grid indices)."""
    F = np.asarray(F, np.float64)
    shape = F.shape[:3]
    if positions is None:
        positions = np.stack(np.indices(shape, dtype=np.float64), axis=-1)
    u = F - positions
    Cx, Cy, Cz = (dct_rows(n, ncoef) for n in shape)
    out = np.empty((3, ncoef, ncoef, ncoef))
    for n in range(3):
        t = u[..., n] @ Cz.T                       # [X, Y, c]
        t = np.einsum("bj,xjc->xbc", Cy, t)        # [X, b, c]
        out[n] = np.einsum("ai,ibc->abc", Cx, t)   # [a, b, c]
    return out / np.sqrt(float(np.prod(shape)))


def coefficients_field(coef, shape, resolution=1):
    """Host evaluation of the warm-start field of a pyramid level
    (Registration.set_field_from_dct): [ceil(N/res)..., 3] positions."""
    coef = np.asarray(coef, np.float64)
    C = int(round((coef.size / 3) ** (1.0 / 3.0)))
    coef = coef.reshape(3, C, C, C)
    shape = tuple(int(n) for n in shape)
    r = int(resolution)
    lvl = tuple(-(-n // r) for n in shape)
    Cs = [dct_rows(n, C)[:, ::r] for n in shape]  # basis at full indices r * p
    amp = np.sqrt(float(np.prod(shape)))
    out = np.empty(lvl + (3,))
    grid = np.indices(lvl, dtype=np.float64)
    for n in range(3):
        u = np.einsum("abc,ai,bj,ck->ijk", coef[n], Cs[0], Cs[1], Cs[2], optimize=True)
        out[..., n] = grid[n] + amp * u / r
    return out


class WarmStartNet:
    """The warm-start network on the GPU (handle over the C ABI)."""

    def __init__(self, grid, max_batch=8, seed=0, ncoef=NCOEF):
        self._api = api = _Api.get()
        self._h = None
        self.grid = tuple(int(n) for n in grid)
        self.max_batch = int(max_batch)
        self.ncoef = int(ncoef)
        h = ctypes.c_void_p()
        api["check"](api["create"](ctypes.byref(h), *self.grid, self.max_batch, self.ncoef,
                                   int(seed)))
        self._h = h
        n = ctypes.c_longlong()
        api["check"](api["nparam"](self._h, ctypes.byref(n)))
        self.nparam = int(n.value)

    def __del__(self):
        if getattr(self, "_h", None):
            self._api["destroy"](self._h)
            self._h = None

    def _call(self, name, *args):
        self._api["check"](self._api[name](self._h, *args))

    def _p(self, a):
        return self._api["ptr"](a)

    def get_params(self):
        out = np.empty(self.nparam, np.float32)
        self._call("get", self._p(out))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters")
        self._call("set", self._p(p))

    def set_dataset(self, fixed_bins, moving_bins, targets=None, L=32, K=32):
        """fixed_bins / moving_bins: [n, *grid] uint8; targets [n, 3, C, C, C]."""
        f = np.ascontiguousarray(fixed_bins, np.uint8)
        m = np.ascontiguousarray(moving_bins, np.uint8)
        n = f.shape[0]
        if f.shape != (n,) + self.grid or m.shape != f.shape:
            raise ValueError(f"bins must be [n, {self.grid}]")
        t = None
        if targets is not None:
            t = np.ascontiguousarray(targets, np.float32)
            if t.size != n * 3 * self.ncoef ** 3:
                raise ValueError(f"targets must be [n, 3, {self.ncoef}, {self.ncoef}, "
                                 f"{self.ncoef}]")
        self._call("data", n, self._p(f), self._p(m), None if t is None else self._p(t),
                   int(L), int(K))

    def train(self, idx, lr=1e-3):
        """Adam steps over idx [steps, batch] (sample indices); returns the
        loss of each step (before its update)."""
        ix = np.ascontiguousarray(idx, np.int32)
        steps, batch = ix.shape
        losses = np.empty(steps, np.float64)
        self._call("train", self._p(ix), steps, batch, float(lr), self._p(losses))
        return losses

    def gradients(self, idx):
        ix = np.ascontiguousarray(idx, np.int32).ravel()
        g = np.empty(self.nparam, np.float32)
        loss = ctypes.c_double()
        self._call("grad", self._p(ix), ix.size, ctypes.byref(loss), self._p(g))
        return float(loss.value), g

    def predict_dataset(self, idx):
        ix = np.ascontiguousarray(idx, np.int32).ravel()
        C = self.ncoef
        out = np.empty((ix.size, 3, C, C, C), np.float32)
        for s in range(0, ix.size, self.max_batch):
            part = np.ascontiguousarray(ix[s:s + self.max_batch])
            self._call("pred_data", self._p(part), part.size, self._p(out[s:s + part.size]))
        return out

    def predict(self, fixed_bins, moving_bins, L=32, K=32):
        """Coefficients [B, 3, C, C, C] for bins [B, *grid] (or one pair [*grid])."""
        f = np.ascontiguousarray(fixed_bins, np.uint8)
        m = np.ascontiguousarray(moving_bins, np.uint8)
        single = f.ndim == 3
        if single:
            f, m = f[None], m[None]
        B = f.shape[0]
        if f.shape[1:] != self.grid or m.shape != f.shape or B > self.max_batch:
            raise ValueError(f"bins must be [<= {self.max_batch}, {self.grid}]")
        C = self.ncoef
        out = np.empty((B, 3, C, C, C), np.float32)
        self._call("pred", B, self._p(f), self._p(m), int(L), int(K), self._p(out))
        return out[0] if single else out

    def predict_pair(self, full_binned_voxels, full_binned_nodes, L=32, K=32):
        """Coefficients for one pair from its full-resolution bins (in
        [0, L) and [0, K); L, K <= 256)."""
        if not (2 <= L <= 256 and 2 <= K <= 256):
            raise ValueError("bin counts must be in [2, 256]")
        f, m = pair_inputs(full_binned_voxels, full_binned_nodes, self.grid, check=False)
        return self.predict(f, m, L, K)

    def save(self, path, **meta):
        np.savez(path, params=self.get_params(), grid=np.array(self.grid),
                 ncoef=np.array(self.ncoef), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, max_batch=1):
        with np.load(path, allow_pickle=False) as z:
            # files written before the count was stored hold 8 x 8 x 8 networks
            ncoef = int(z["ncoef"]) if "ncoef" in z.files else 8
            net = cls(tuple(int(n) for n in z["grid"]), max_batch=max_batch, ncoef=ncoef)
            net.set_params(z["params"])
        return net


WEIGHTS_C2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                          "warmstart_160x192x160.npz")
