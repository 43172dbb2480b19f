"""Seeded synthetic registration pairs for the benchmark and test workloads.

Fixed image: a phantom of random ellipses (2D) / ellipsoids (3D) painted over
a zero background, each region one of four tissue intensities
{0.1, 0.35, 0.6, 0.85}, Gaussian-smoothed (sigma 1 voxel) with additive
N(0, 0.02) noise.  Moving image: the fixed image resampled through a smooth
random displacement whose lowest 8 (per axis) orthonormal-DCT coefficients per
component are N(0,1), scaled to a target RMS displacement, then passed
through the monotone map x -> x^0.7 (so that a histogram likelihood, not
intensity equality, is needed), plus fresh noise.

The registration inputs follow BINDER/FastRegistration.py:206-225 and
:368-426: robust 32-bin binning, theta = 1/L, identity voxel grids, and
gamma_n = gamma * prod(voxel size) * Nv * node_size_n^2.
"""
from dataclasses import dataclass, field

import numpy as np
from scipy import ndimage
from scipy.fft import idctn

from .utils import bin_signal_robust, initialize_fourier_domain_smoothing

TISSUE = (0.1, 0.35, 0.6, 0.85)


def phantom(shape, n_shapes, rng, semi=(10.0, 60.0), sigma=1.0, noise=0.02):
    shape = tuple(int(s) for s in shape)
    img = np.zeros(shape, np.float64)
    lo, hi = semi
    for _ in range(n_shapes):
        c = [rng.uniform(0.15 * s, 0.85 * s) for s in shape]
        r = [min(rng.uniform(lo, hi), 0.45 * s) for s in shape]
        # paint inside the bounding box only: outside it some axis term
        # ((g - c) / r)^2 alone exceeds 1
        box = tuple(slice(max(0, int(np.floor(ci - ri))), min(s, int(np.ceil(ci + ri)) + 1))
                    for ci, ri, s in zip(c, r, shape))
        axes = np.ix_(*[np.arange(b.start, b.stop, dtype=np.float64) for b in box])
        m = sum(((g - ci) / ri) ** 2 for g, ci, ri in zip(axes, c, r)) <= 1.0
        img[box][m] = TISSUE[rng.integers(0, len(TISSUE))]
    img = ndimage.gaussian_filter(img, sigma)
    return img + rng.normal(0.0, noise, shape)


def _dct_rows(n, k):
    """Rows a < k of the orthonormal DCT-II matrix of length n, [k][n]."""
    a = np.arange(k)[:, None]
    i = np.arange(n)[None, :]
    C = np.sqrt(2.0 / n) * np.cos(np.pi * (2 * i + 1) * a / (2 * n))
    C[0] = np.sqrt(1.0 / n)
    return C


def dct_displacement(shape, rms, rng, ncoef=8, separable=False):
    """[*shape, D] smooth displacement, lowest ncoef^D DCT coefficients N(0,1).
    separable=True synthesises the field from the ncoef basis rows per axis
    (the same field up to rounding, several times faster at 160x192x160)."""
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    out = np.empty(shape + (D,), np.float64)
    for n in range(D):
        sl = tuple(slice(0, min(ncoef, s)) for s in shape)
        coef = rng.standard_normal(tuple(x.stop for x in sl))
        if separable:
            f = coef
            for ax, s in enumerate(shape):
                C = _dct_rows(s, coef.shape[ax])  # [k][s]
                f = np.moveaxis(np.tensordot(C.T, np.moveaxis(f, ax, 0), axes=1), 0, ax)
            out[..., n] = f
        else:
            c = np.zeros(shape)
            c[sl] = coef
            out[..., n] = idctn(c, norm="ortho")
    cur = np.sqrt(np.mean(np.sum(out ** 2, axis=-1)))
    return out * (rms / max(cur, 1e-12))


@dataclass
class Pair:
    fixed: np.ndarray
    moving: np.ndarray
    displacement: np.ndarray
    binned_fixed: np.ndarray = field(default=None)
    binned_moving: np.ndarray = field(default=None)


def make_pair(shape, n_shapes, rms, seed, semi=(10.0, 60.0), bins=32, gamma_map=0.7,
              separable=False):
    rng = np.random.default_rng(seed)
    fixed = phantom(shape, n_shapes, rng, semi=semi)
    u = dct_displacement(shape, rms, rng, separable=separable)
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    coords = [g + u[..., a] for a, g in enumerate(grids)]
    warped = ndimage.map_coordinates(fixed, coords, order=1, mode="nearest")
    moving = np.clip(warped, 0.0, None) ** gamma_map + rng.normal(0.0, 0.02, warped.shape)
    p = Pair(fixed=fixed, moving=moving, displacement=u)
    p.binned_fixed, _ = bin_signal_robust(fixed, bins)
    p.binned_moving, _ = bin_signal_robust(moving, bins)
    return p


# FastRegistration keywords of the 3D workloads (run_BINDER's defaults for
# this path: 32 bins, gamma 1e-6 per voxel, gamma raised at coarse levels,
# topology correction, convergence threshold 2.5e-5), pyramid [8, 4, 2, 1]
# with a 1000 x resolution iteration cap per level
REGISTRATION_KW = dict(num_voxel_intesity_levels=32, num_node_intensity_levels=32, gamma=1e-6,
                       topology_correction=True, increase_gamma_at_DS_resolution=True,
                       alpha=2.0, convergence_th=2.5e-5)
PYRAMID = (8, 4, 2, 1)
EM_ITERATIONS = 1000

CONFIGS = {
    # 2D MI registration, 192x160, 6 ellipses, 3 px RMS warp
    "c1": dict(shape=(192, 160), n_shapes=6, rms=3.0),
    # 3D MI registration + topology correction + sampler, 160x192x160
    "c2": dict(shape=(160, 192, 160), n_shapes=12, rms=6.0),
    # precision variant, 256x256x192
    "c3": dict(shape=(256, 256, 192), n_shapes=12, rms=6.0),
}


@dataclass
class Problem:
    """Everything the _core.Registration calls need, for one resolution."""
    shape: tuple
    final_locations: np.ndarray  # [Nx,Ny,Nz,N] float64
    voxel_pos: np.ndarray
    binned_nodes: np.ndarray     # int32 [Nx,Ny,Nz]
    binned_voxels: np.ndarray
    theta: np.ndarray            # [K,L]
    D: np.ndarray
    gammas: tuple
    K: int
    L: int
    alpha: float
    nnz: int


def problem_from_pair(pair, gamma=1e-6, alpha=2.0, bins=32, fd_order=2):
    """Registration inputs at full resolution with identity initialisation
    (BINDER/FastRegistration.py:206-225, :368-426, :411-417)."""
    fixed_b = np.ascontiguousarray(pair.binned_fixed, np.int32)
    moving_b = np.ascontiguousarray(pair.binned_moving, np.int32)
    if fixed_b.ndim == 2:
        fixed_b = fixed_b[:, :, None]
        moving_b = moving_b[:, :, None]
    shp = fixed_b.shape
    is2d = shp[2] == 1
    N = 2 if is2d else 3
    pos = np.stack(np.indices(shp, dtype=np.float64), axis=-1)[..., :N]
    nv = int(np.prod(shp))
    g = gamma * 1.0 * nv * 1.0
    D = initialize_fourier_domain_smoothing(list(shp[:N]), True, fd_order, N,
                                            voxel_size=np.ones(3))
    D = np.ascontiguousarray(D.reshape(shp))
    theta = np.ones((bins, bins)) / bins
    return Problem(shape=shp, final_locations=pos.copy(), voxel_pos=pos, binned_nodes=moving_b,
                   binned_voxels=fixed_b, theta=theta, D=D, gammas=(g, g, g), K=bins, L=bins,
                   alpha=alpha, nnz=int(np.count_nonzero(fixed_b)))


def build_registration(core, prob, threads=1, spline_order=3):
    """Construct a `core.Registration` (binder_amd._core or the oracle's
    mirror) for `prob` with the MI likelihood and the DCT prior set, its node
    weights B-splines of `spline_order` (1, 2 or 3)."""
    Nx, Ny, Nz = prob.shape
    r = core.Registration(prob.final_locations.ravel(), threads, prob.voxel_pos.ravel(),
                          prob.voxel_pos.ravel(), Nx, Ny, Nz, Nx, Ny, Nz, spline_order,
                          prob.nnz)
    r.setNonLinearTransformation(*prob.gammas, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, prob.D.ravel())
    r.setMILikelihood(prob.alpha, prob.theta.ravel(), prob.K, prob.L, threads,
                      prob.binned_nodes.ravel(), prob.binned_voxels.ravel())
    return r


def config_problem(name, seed=0, scale=None, **kw):
    """Problem for a named config; `scale` shrinks the grid (tests)."""
    cfg = dict(CONFIGS[name])
    shape = cfg["shape"]
    semi = (10.0, 60.0)
    if scale is not None:
        shape = tuple(max(4, int(round(s * scale))) for s in shape)
        semi = (max(1.5, 10.0 * scale), max(3.0, 60.0 * scale))
    pair = make_pair(shape, cfg["n_shapes"], cfg["rms"] * (scale or 1.0), seed, semi=semi)
    return pair, problem_from_pair(pair, **kw)
