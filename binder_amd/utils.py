"""Host-side producers of the path's inputs (NumPy), restated from
BINDER/utils.py: intensity binning and the DCT-domain eigenvalue array D."""
import numpy as np


def bin_signal(signal, number_of_bins):
    """Uniform binning to rounded bin indices (BINDER/utils.py:12-16)."""
    signal = np.asarray(signal, dtype=np.float64)
    lo = np.min(signal)
    width = (np.max(signal) - lo) / (number_of_bins - 1)
    return np.round((signal - lo) / width).astype(np.int32)


def bin_signal_robust(signal, number_of_bins, lower_percentile=0.01, higher_percentile=0.99,
                      intensity_bins=None):
    """Robust binning between the 1st/99th percentiles of a 1000-bin histogram
    (BINDER/utils.py:20-82).  Returns (binned int32, bin edges)."""
    signal = np.asarray(signal, dtype=np.float64)
    if intensity_bins is not None:
        edges = np.asarray(intensity_bins)
        b = np.digitize(signal, edges).astype(np.int32) - 1
        b[signal <= edges[0]] = 0
        b[signal >= edges[-1]] = number_of_bins - 1
        return b, edges
    core = _host_core()
    if core is not None:
        hist, centers = core.histogram_uniform(signal, 1000)
    else:
        hist, centers = np.histogram(signal, 1000)
    cdf = np.cumsum(hist / np.sum(hist))

    def _edge(p):
        idx = np.where(cdf <= p)[0]
        add = 0.01
        while len(idx) == 0:
            idx = np.where(cdf <= p + add)[0]
            add += 0.01
        return centers[idx[-1]]

    lo = _edge(lower_percentile)
    hi = _edge(higher_percentile)
    width = (hi - lo) / (number_of_bins - 1)
    edges = np.linspace(lo, hi, number_of_bins + 1)
    if core is not None:
        return core.bin_floor(signal, lo, hi, width, number_of_bins), edges
    b = np.floor((signal - lo) / width).astype(np.int32)
    b[signal < lo] = 0
    b[signal > hi] = number_of_bins - 1
    return b, edges


def _host_core():
    """The library's multithreaded host binning (numpy's arithmetic), or None
    where the library is not built (numpy then does the same computation)."""
    try:
        from . import _core
        return _core
    except ImportError:
        return None


def initialize_fourier_domain_smoothing(size, DCT=True, finite_difference_order=2,
                                        number_of_dimensions=3, return_full_matrix=False,
                                        voxel_size=np.ones(3)):
    """Eigenvalues D of the finite-difference prior in the DCT (or real-FFT)
    domain: D = (sum_a (2 - 2 cos(pi k_a / N_a)) / vs_a^2) ** order
    (BINDER/utils.py:270-365)."""
    vs = np.asarray(voxel_size, dtype=np.float64)
    order = finite_difference_order
    if number_of_dimensions == 1:
        n = size[0]
        if DCT:
            k = np.arange(n * 2)
            D = ((2.0 - 2 * np.cos(np.pi * k / n)) / vs[0] ** 2) ** order
            return (D[:n], D) if return_full_matrix else D[:n]
        k = np.arange(n)
        D = ((2.0 - 2 * np.cos(2 * np.pi * k / n)) / vs[0] ** 2) ** order
        half = int(np.floor(len(D) / 2)) + 1
        return (D[:half], D) if return_full_matrix else D[:half]
    dims = list(size[:number_of_dimensions])
    nd = len(dims)
    if DCT:
        # separable: the per-axis terms are 1-D and broadcast (same values and
        # the same x + y + z summation order as evaluating on a full grid)
        D = None
        for a, n in enumerate(dims):
            k = np.arange(2 * n if return_full_matrix else n, dtype=np.float64)
            t = (2.0 - 2 * np.cos(np.pi * k / n)) / vs[a] ** 2
            t = t.reshape([-1 if b == a else 1 for b in range(nd)])
            D = t if D is None else D + t
        D = D ** order
        if return_full_matrix:
            return D[tuple(slice(0, n) for n in dims)], D
        return D
    grids = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    D = sum((2.0 - 2 * np.cos(2 * np.pi * g / n)) / vs[a] ** 2
            for a, (g, n) in enumerate(zip(grids, dims)))
    D = D ** order
    half = int(np.floor(D.shape[number_of_dimensions - 1] / 2)) + 1
    if return_full_matrix:
        return D[..., :half], D
    return D[..., :half]


def hist_norm(source, template):
    """Histogram matching of `source` to `template` by quantile mapping
    (BINDER/utils.py:84-100)."""
    src = np.asarray(source)
    s_vals, s_idx, s_cnt = np.unique(src.ravel(), return_inverse=True, return_counts=True)
    t_vals, t_cnt = np.unique(np.asarray(template).ravel(), return_counts=True)
    s_q = np.cumsum(s_cnt).astype(np.float64)
    s_q /= s_q[-1]
    t_q = np.cumsum(t_cnt).astype(np.float64)
    t_q /= t_q[-1]
    mapped = np.interp(s_q, t_q, t_vals).astype(src.dtype)
    return mapped[s_idx].reshape(src.shape)


def maskImage(image, th_value=0, use_robust_minimum=False):
    """Foreground mask: threshold (optionally at the 10th percentile of a
    1000-bin histogram), one binary closing, hole filling
    (BINDER/utils.py:242-267).  int32."""
    from scipy import ndimage
    image = np.asarray(image)
    if use_robust_minimum:
        hist, centers = np.histogram(image.ravel(), 1000)
        cdf = np.cumsum(hist / np.sum(hist))
        p = 0.1
        idx = np.where(cdf <= p)[0]
        while len(idx) == 0:
            p += 0.01
            idx = np.where(cdf <= p)[0]
        th_value = centers[idx[-1]]
    mask = ndimage.binary_closing(image > th_value, iterations=1)
    mask = ndimage.binary_fill_holes(mask)
    return mask.astype(np.int32)


def save_warped_image(deformed_position, movingImage, outputDir, order=1, affine=np.eye(4),
                      outputName="warped.nii.gz"):
    """Moving image resampled at the deformed positions, saved as NIfTI
    (BINDER/utils.py:222-227)."""
    import os
    from . import nifti, resample
    pos = np.asarray(deformed_position, dtype=np.float64)
    img = resample.warp(movingImage, np.moveaxis(pos, -1, 0), order=order)
    nifti.save(nifti.NiftiImage(img, affine), os.path.join(outputDir, outputName))
    print("Warped image saved")


def save_warp(deformations, outputDir, affine=np.eye(4), outputName="warp.nii.gz"):
    """BINDER/utils.py:230-233."""
    import os
    from . import nifti
    nifti.save(nifti.NiftiImage(np.asarray(deformations), affine),
               os.path.join(outputDir, outputName))
    print("Warp saved")


def save_image(generalImage, outputDir, outputName, affine=np.eye(4)):
    """BINDER/utils.py:236-239."""
    import os
    from . import nifti
    nifti.save(nifti.NiftiImage(np.asarray(generalImage), affine),
               os.path.join(outputDir, outputName))
    print("Image saved")


def save_likelihood_parameters(registrator, outputDir):
    """MI likelihood parameters to likelihood_parameters.npz
    (BINDER/utils.py:443-472, MI branch)."""
    import os
    np.savez(os.path.join(outputDir, "likelihood_parameters.npz"), metric=registrator.metric,
             theta=registrator.theta, alpha=registrator.alpha)
