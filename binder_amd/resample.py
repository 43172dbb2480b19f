"""Host-side resampling used around the registration path, written from the
published behaviour of ``skimage.transform.resize`` and
``skimage.transform.warp`` (scikit-image 0.19+), which the reference calls
(BINDER/FastRegistration.py:4, :566-568; BINDER/utils.py:7, :222-227;
BINDER/cli/run_BINDER.py:2, :228-236) and which is not available here.

resize(image, shape, order):
  * trailing dimensions not named in ``shape`` are kept;
  * when any axis shrinks, a Gaussian anti-aliasing filter with
    sigma_a = max(0, (in_a / out_a - 1) / 2) is applied first;
  * interpolation with scipy.ndimage.zoom on the full pixel extent
    (grid_mode=True), factor out_a / in_a per axis;
  * boundary mode "reflect" of scikit-image, which is scipy's "mirror";
  * the output is clipped to the input's [min, max].

warp(image, coords, order): scipy.ndimage.map_coordinates at the given
coordinate map, constant 0 outside (scikit-image's "constant" mode on the
pixel grid, scipy's "grid-constant"), clipped to the input's range, values
kept in the input's units (``preserve_range=True``).

The exact agreement with scikit-image is not checked here (scikit-image is
absent from this environment): constant fields, the identity and the
clipping are tested instead.
"""
import numpy as np
from scipy import ndimage


def resize(image, output_shape, order=1):
    img = np.asarray(image, dtype=np.float64)
    out_shape = tuple(int(s) for s in output_shape)
    out_shape = out_shape + img.shape[len(out_shape):]
    if len(out_shape) != img.ndim:
        raise ValueError("output_shape has more dimensions than the image")
    factors = np.array(img.shape, dtype=np.float64) / np.array(out_shape, dtype=np.float64)
    lo, hi = img.min(), img.max()
    if np.any(factors > 1.0) and order > 0:
        sigma = np.maximum(0.0, (factors - 1.0) / 2.0)
        img = ndimage.gaussian_filter(img, sigma, mode="mirror")
    zoom = np.array(out_shape, dtype=np.float64) / np.array(img.shape, dtype=np.float64)
    if np.all(zoom == 1.0):
        out = img.copy()
    else:
        out = ndimage.zoom(img, zoom, order=order, mode="mirror", grid_mode=True)
    if out.shape != out_shape:  # zoom rounds the shape; never expected
        raise RuntimeError(f"resize produced {out.shape}, wanted {out_shape}")
    np.clip(out, lo, hi, out=out)
    return out


def warp(image, coords, order=1, output_shape=None):
    """Sample `image` at `coords` (shape [ndim_image, *output]) as
    skimage.transform.warp(image, coords, order=order, preserve_range=True)."""
    img = np.asarray(image, dtype=np.float64)
    c = np.asarray(coords, dtype=np.float64)
    # with a coordinate map the output grid is the map's own (the reference
    # passes the field's shape as output_shape, which the map overrides)
    del output_shape
    if c.shape[0] != img.ndim:
        raise ValueError(f"coordinate map has {c.shape[0]} rows for a {img.ndim}-D image")
    out = ndimage.map_coordinates(img, c, order=order, mode="grid-constant", cval=0.0,
                                  prefilter=order > 1)
    lo, hi = min(img.min(), 0.0), max(img.max(), 0.0)
    np.clip(out, lo, hi, out=out)
    return out
