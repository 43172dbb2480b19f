"""CPU checks of the warm-start network's FP64 oracle (oracle/cnn.py) and of
the host-side coefficient helpers (binder_amd/warmstart.py).

The network has no counterpart in the reference, so the oracle is pinned by
library routines and identities: scipy.ndimage.correlate for the
convolution, the adjoint identity for pooling / up-sampling, scipy's
orthonormal dctn for the head, and central finite differences for the whole
backward pass.
"""
import numpy as np
import pytest
from scipy import ndimage
from scipy.fft import dctn, idctn

import cnn


def test_conv3d_matches_scipy_correlate(rng):
    x = rng.standard_normal((1, 6, 5, 7, 3))
    W = rng.standard_normal((2, 27, 3))
    b = rng.standard_normal(2)
    y = cnn.conv3d(x, W, b)
    for co in range(2):
        ref = np.full(x.shape[1:4], b[co])
        for ci in range(3):
            k = W[co, :, ci].reshape(3, 3, 3)  # tap t = 9(dx+1) + 3(dy+1) + (dz+1)
            ref += ndimage.correlate(x[0, ..., ci], k, mode="constant", cval=0.0)
        np.testing.assert_allclose(y[0, ..., co], ref, rtol=1e-12, atol=1e-12)


def test_conv3d_backward_is_the_adjoint(rng):
    x = rng.standard_normal((2, 4, 6, 4, 3))
    W = rng.standard_normal((5, 27, 3))
    g = rng.standard_normal((2, 4, 6, 4, 5))
    dx, dW, db = cnn.conv3d_backward(x, W, g)
    # <conv(x), g> is linear in x and in W
    lhs = np.sum(cnn.conv3d(x, W, np.zeros(5)) * g)
    assert abs(lhs - np.sum(dx * x)) < 1e-9 * abs(lhs) + 1e-9
    assert abs(lhs - np.sum(dW * W)) < 1e-9 * abs(lhs) + 1e-9
    np.testing.assert_allclose(db, g.sum(axis=(0, 1, 2, 3)))


def test_pool_and_up_adjoints(rng):
    x = rng.standard_normal((2, 4, 6, 8, 3))
    y = rng.standard_normal((2, 2, 3, 4, 3))
    assert np.isclose(np.sum(cnn.pool(x) * y), np.sum(x * cnn.pool_backward(y)))
    assert np.isclose(np.sum(cnn.up(y) * x), np.sum(y * cnn.up_backward(x)))
    # pooling a constant keeps it, up-sampling then pooling is the identity
    np.testing.assert_allclose(cnn.pool(cnn.up(y)), y)


@pytest.mark.parametrize("C", [8, 16])
def test_head_projection_is_the_truncated_orthonormal_dct(rng, C):
    """C = 16 exceeds the grid (10 x 12 x 9): the coefficients it does not
    have are zero."""
    U = rng.standard_normal((2, 3, 10, 12, 9))
    c = cnn.head_project(U, C)
    assert c.shape == (2, 3, C, C, C)
    for b in range(2):
        for n in range(3):
            ref = np.zeros((C, C, C))
            full = dctn(U[b, n], norm="ortho")[:C, :C, :C]
            ref[:full.shape[0], :full.shape[1], :full.shape[2]] = full
            ref /= np.sqrt(10 * 12 * 9)
            np.testing.assert_allclose(c[b, n], ref, rtol=1e-10, atol=1e-12)
    # the expansion is its adjoint
    dc = rng.standard_normal(c.shape)
    assert np.isclose(np.sum(c * dc), np.sum(U * cnn.head_expand(dc, U.shape[2:])))


def test_round_bf16_ties_to_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0])
    out = cnn.round_bf16(vals)
    np.testing.assert_array_equal(out, [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 0.0])


def _small_problem(rng, grid=(32, 32, 32), batch=1):
    fb = rng.integers(0, 32, (batch,) + grid)
    mb = rng.integers(0, 32, (batch,) + grid)
    x0 = cnn.make_input(fb, mb, bf16=False)
    p = np.zeros(cnn.num_params())
    off = 0
    for ci, co in cnn.CONV_LAYERS:
        fan = 27 * (2 if ci == 8 else ci)
        w = rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / fan)
        if ci == 8:
            w[:, :, 2:] = 0.0
        p[off:off + w.size] = w.ravel()
        off += w.size
        p[off:off + co] = 0.05 * rng.standard_normal(co)
        off += co
    p[off:] = 0.1 * rng.standard_normal(195)
    t = 0.1 * rng.standard_normal((batch, 3, 8, 8, 8))
    return x0, p, t


def test_backward_matches_finite_differences(rng):
    x0, p, t = _small_problem(rng)

    def loss(pp):
        net = cnn.Net(pp, bf16=False)
        return cnn.mse(net.forward(x0), t)[0]

    net = cnn.Net(p, bf16=False)
    L, dc = cnn.mse(net.forward(x0), t)
    g = net.backward(dc)
    assert g.shape == p.shape
    # one weight and one bias of every conv layer, plus head entries
    picks, off = [], 0
    for ci, co in cnn.CONV_LAYERS:
        nw = co * 27 * ci
        picks.append(off + 13 * ci + 1 if ci != 8 else off + 13 * ci)  # centre tap
        picks.append(off + nw + 1)
        off += nw + co
    picks += [off + 5, off + 192 + 2]
    h = 1e-5
    for i in picks:
        e = np.zeros_like(p)
        e[i] = h
        fd = (loss(p + e) - loss(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3), (i, fd, g[i])


@pytest.mark.parametrize("C", [8, 16])
def test_field_coefficients_round_trip(rng, C):
    from binder_amd.warmstart import coefficients_field, field_coefficients
    shape = (20, 24, 18)
    coef = rng.standard_normal((3, C, C, C)) * 0.1
    F = coefficients_field(coef, shape)
    np.testing.assert_allclose(field_coefficients(F, ncoef=C), coef, atol=1e-12)
    # the same field through scipy: zero-padded coefficients, idctn
    for n in range(3):
        c = np.zeros(shape)
        c[:C, :C, :C] = coef[n] * np.sqrt(np.prod(shape))
        np.testing.assert_allclose(F[..., n] - np.indices(shape)[n], idctn(c, norm="ortho"),
                                   atol=1e-12)
    # a pyramid level samples the full-resolution displacement at 2p, / 2
    F2 = coefficients_field(coef, shape, resolution=2)
    assert F2.shape == (10, 12, 9, 3)
    np.testing.assert_allclose(F2 - np.stack(np.indices((10, 12, 9)), -1),
                               (F - np.stack(np.indices(shape), -1))[::2, ::2, ::2] / 2,
                               atol=1e-12)


def test_network_grid_and_inputs():
    from binder_amd.warmstart import network_grid, num_params, pair_inputs
    assert network_grid((160, 192, 160)) == (80, 96, 80)
    assert network_grid((96, 112, 129)) == (48, 56, 72)
    assert network_grid((40, 30, 20)) == (32, 32, 32)
    fb = np.arange(9 * 7 * 5).reshape(9, 7, 5) % 32
    f, m = pair_inputs(fb, fb, (32, 32, 32))
    assert f.dtype == np.uint8 and f.shape == (32, 32, 32)
    np.testing.assert_array_equal(f[:5, :4, :3], fb[::2, ::2, ::2])
    assert not f[5:].any() and not f[:, 4:].any() and not f[:, :, 3:].any()
    assert num_params() == cnn.num_params()
    with pytest.raises(ValueError):
        pair_inputs(fb + 300, fb, (32, 32, 32))
