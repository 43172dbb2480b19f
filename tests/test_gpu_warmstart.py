"""Warm-start network on the GPU against its FP64 oracle (oracle/cnn.py).

The oracle rounds the tensors the HIP path stores in BF16 the same way, so
the comparison measures only accumulation order (FP32 MFMA against FP64):
forward coefficients within 1 % relative L2; every layer's parameter
gradient within 1.5 x the difference that BF16 storage itself makes (the
BF16 oracle against the unrounded one, 1-17 % at 32 x 40 x 32) and at
cosine similarity above 0.98.
"""
import numpy as np
import pytest

import cnn

pytestmark = pytest.mark.gpu

GRID = (32, 40, 32)


@pytest.fixture(scope="module")
def ws():
    from binder_amd import warmstart
    return warmstart


def _data(rng, n, grid=GRID, C=8):
    # smooth blobs rather than white noise, so ReLUs are not all on a knife edge
    from scipy import ndimage
    fb = np.empty((n,) + grid, np.uint8)
    mb = np.empty((n,) + grid, np.uint8)
    for i in range(n):
        for a in (fb, mb):
            f = ndimage.gaussian_filter(rng.standard_normal(grid), 2.0)
            f = (f - f.min()) / (f.max() - f.min())
            a[i] = np.minimum((f * 32).astype(np.int64), 31)
    t = (0.2 * rng.standard_normal((n, 3, C, C, C))).astype(np.float32)
    return fb, mb, t


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(b), 1e-30))


def _layer_slices():
    out, off = [], 0
    for i, (ci, co) in enumerate(cnn.CONV_LAYERS):
        out.append((f"conv{i}.W", slice(off, off + co * 27 * ci)))
        off += co * 27 * ci
        out.append((f"conv{i}.b", slice(off, off + co)))
        off += co
    out.append(("head.W", slice(off, off + 192)))
    out.append(("head.b", slice(off + 192, off + 195)))
    return out


@pytest.mark.parametrize("C", [8, 16])
def test_forward_matches_oracle(ws, rng, C):
    """C = 16 on this grid (head grid 8 x 10 x 8) also checks the zero
    coefficients at frequencies the head grid does not have."""
    print(f"warm-start forward, grid {GRID}, batch 2, {C} coefficients", flush=True)
    net = ws.WarmStartNet(GRID, max_batch=2, seed=3, ncoef=C)
    p = net.get_params()
    # a head large enough for the coefficients to carry signal
    p[-195:] = (0.05 * rng.standard_normal(195)).astype(np.float32)
    net.set_params(p)
    fb, mb, _ = _data(rng, 2)
    got = net.predict(fb, mb)
    ref = cnn.Net(p, bf16=True, ncoef=C).forward(cnn.make_input(fb, mb))
    assert got.shape == (2, 3, C, C, C)
    assert _rel(got, ref) < 1e-2, _rel(got, ref)


@pytest.mark.parametrize("C", [8, 16])
def test_gradients_match_oracle(ws, rng, C):
    print(f"warm-start gradients, grid {GRID}, batch 2, {C} coefficients", flush=True)
    net = ws.WarmStartNet(GRID, max_batch=2, seed=5, ncoef=C)
    p = net.get_params()
    p[-195:] = (0.05 * rng.standard_normal(195)).astype(np.float32)
    net.set_params(p)
    fb, mb, t = _data(rng, 3, C=C)
    net.set_dataset(fb, mb, t)
    idx = np.array([2, 0], np.int32)
    loss, g = net.gradients(idx)
    o = cnn.Net(p, bf16=True, ncoef=C)
    c = o.forward(cnn.make_input(fb[idx], mb[idx]))
    lo, dc = cnn.mse(c, t[idx])
    go = o.backward(dc)
    assert abs(loss - lo) <= 1e-2 * lo
    # the same gradient without any BF16 rounding: the scale of the
    # differences that BF16 storage itself causes
    o64 = cnn.Net(p, bf16=False, ncoef=C)
    c64 = o64.forward(cnn.make_input(fb[idx], mb[idx], bf16=False))
    g64 = o64.backward(cnn.mse(c64, t[idx])[1])
    worst = []
    for name, sl in _layer_slices():
        a, b, e = g[sl], go[sl], g64[sl]
        if name == "conv0.W":  # only input channels 0 and 1 carry data
            assert not np.any(a.reshape(16, 27, 8)[:, :, 2:])
            a, b, e = (v.reshape(16, 27, 8)[:, :, :2] for v in (a, b, e))
        a, b, e = np.ravel(a), np.ravel(b), np.ravel(e)
        cos = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))
        print(f"{name:10s} rel {_rel(a, b):.2e} cos {cos:.5f}  (bf16 vs fp64 oracle "
              f"{_rel(b, e):.2e})", flush=True)
        worst.append((_rel(a, b), cos, _rel(b, e), name))
    # the HIP gradient sits closer to the BF16 oracle than BF16 storage
    # itself moves the gradient (a wrong tap, mask or transpose gives a
    # relative error near 1 and a cosine near 0).  With C = 16 on this grid
    # most targets are unreachable coefficients, which the HIP and the
    # oracle's BF16 roundings then weight differently in the convolution
    # layers: those are held to the direction (the same kernels as C = 8),
    # the head, where the count enters, to the full bound.
    for r, cos, rb, name in worst:
        if C == 8 or name.startswith("head"):
            assert r < max(3e-2, 1.5 * rb) and cos > 0.98, (name, r, cos, rb)
        else:
            assert cos > 0.99, (name, r, cos, rb)


def test_predict_paths_agree_adam_step_and_training(ws, rng):
    print(f"warm-start training, grid {GRID}, 6 samples", flush=True)
    net = ws.WarmStartNet(GRID, max_batch=3, seed=7, ncoef=8)
    fb, mb, t = _data(rng, 6)
    net.set_dataset(fb, mb, t)
    a = net.predict_dataset(np.array([4, 1, 5]))
    b = net.predict(fb[[4, 1, 5]], mb[[4, 1, 5]])
    np.testing.assert_array_equal(a, b)
    # first Adam step: m_hat = g, v_hat = g^2, so p1 = p0 - lr g / (|g| + eps)
    batch = np.array([4, 1, 5], np.int32)
    p0 = net.get_params()
    loss0, g = net.gradients(batch)
    lr = 1e-3
    losses = net.train(batch[None], lr=lr)
    assert losses[0] == loss0
    p1 = net.get_params()
    g64 = g.astype(np.float64)
    expect = p0 - lr * g64 / (np.abs(g64) + 1e-8)
    np.testing.assert_allclose(p1, expect, rtol=1e-5, atol=1e-7)
    # repeated steps on one batch lower its loss
    losses = net.train(np.tile(batch, (40, 1)), lr=lr)
    assert np.all(np.isfinite(losses))
    assert losses[-5:].mean() < losses[:5].mean(), (losses[:5], losses[-5:])
    # parameters round-trip and reproduce the predictions
    p = net.get_params()
    net2 = ws.WarmStartNet(GRID, max_batch=3, seed=99, ncoef=8)
    net2.set_params(p)
    np.testing.assert_array_equal(net2.predict(fb[:3], mb[:3]), net.predict(fb[:3], mb[:3]))


@pytest.mark.parametrize("res,C", [(1, 8), (2, 8), (1, 16), (2, 16)])
def test_field_from_dct_matches_host_evaluation(core, ws, rng, res, C):
    shape = (40, 36, 30)
    print(f"set_field_from_dct, full grid {shape}, resolution {res}, {C} coefficients",
          flush=True)
    lvl = tuple(-(-n // res) for n in shape)
    pos = np.stack(np.indices(lvl, dtype=np.float64), -1)
    r = core.Registration(pos.ravel(), 1, pos.ravel(), pos.ravel(), *lvl, *lvl, 3,
                          int(np.prod(lvl)))
    coef = (0.05 * rng.standard_normal((3, C, C, C))).astype(np.float32)
    r.set_field_from_dct(coef, shape, res)
    got = r.get_final_locations_as_numpy_array()
    ref = ws.coefficients_field(coef, shape, res)
    assert np.max(np.abs(got - ref)) < 2e-4
    with pytest.raises(ValueError):
        r.set_field_from_dct(coef, (shape[0] + 2 * res, shape[1], shape[2]), res)


class _FixedNet:
    """Stands in for the trained network: returns given coefficients (the
    network itself is checked above)."""

    def __init__(self, coef):
        self.coef = np.asarray(coef, np.float32)

    def predict_pair(self, *args, **kwargs):
        return self.coef


def _warm_pyramid(backend_module, monkeypatch, pair, net, levels):
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", backend_module)
    reg = fr.FastRegistration(voxels=pair.fixed, nodes=pair.moving,
                              num_voxel_intesity_levels=32, num_node_intensity_levels=32,
                              gamma=1e-6, convergence_th=2.5e-5, topology_correction=True,
                              increase_gamma_at_DS_resolution=True)
    used = reg.run_warm(net, levels, 10, compare_resolution=4)
    return reg, used


@pytest.mark.parametrize("levels", [(1,), (2, 1)])
def test_warm_pyramid_matches_oracle(core, orc, monkeypatch, ws, levels):
    """The warm-started pipeline (start field evaluated on the GPU) against
    the FP64 oracle running the same pipeline from the host-evaluated start:
    same fallback decision and iteration counts, field within 0.025 voxel
    RMS (the bound of the cold pyramid test)."""
    from binder_amd import synth
    print(f"warm pyramid 48x40x44 levels {levels} (HIP vs oracle)", flush=True)
    pair = synth.make_pair((48, 40, 44), 5, 2.0, seed=3, semi=(3.0, 12.0))
    # start from the low-frequency part of the true inverse warp's estimate
    # by the cold pipeline (what the network is trained to predict)
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", core)
    cold = fr.FastRegistration(voxels=pair.fixed, nodes=pair.moving,
                               num_voxel_intesity_levels=32, num_node_intensity_levels=32,
                               gamma=1e-6, convergence_th=2.5e-5, topology_correction=True,
                               increase_gamma_at_DS_resolution=True)
    cold.run([4, 2, 1], 10)
    net = _FixedNet(ws.field_coefficients(cold.deformedVoxelPositions))
    g, gu = _warm_pyramid(core, monkeypatch, pair, net, levels)
    o, ou = _warm_pyramid(orc, monkeypatch, pair, net, levels)
    assert gu and ou
    assert np.allclose(g.warm_objectives, o.warm_objectives, rtol=1e-5)
    gi = [lv[1] for lv in g.level_log]
    oi = [lv[1] for lv in o.level_log]
    diff = np.asarray(g.deformedVoxelPositions, np.float64) - o.deformedVoxelPositions
    rms = float(np.sqrt(np.mean(np.sum(diff ** 2, -1))))
    print(f"iterations HIP {gi} oracle {oi}; field RMS diff {rms:.2e}", flush=True)
    assert gi == oi
    assert rms < 0.025


def test_warm_start_falls_back_to_the_cold_pyramid(core, monkeypatch, ws):
    """A warm start far from the answer (a large random displacement) has a
    worse start objective than the identity: run_warm then runs the plain
    cold pyramid and returns exactly its result."""
    from binder_amd import synth
    import binder_amd.FastRegistration as fr
    print("warm start fallback 48x40x44", flush=True)
    pair = synth.make_pair((48, 40, 44), 5, 2.0, seed=3, semi=(3.0, 12.0))
    bad = np.zeros((3, 8, 8, 8), np.float32)
    bad[:, 1, 1, 1] = 1.5  # about 1.5 voxel-RMS displacement per component
    bad[:, 2, 0, 3] = -1.5
    w, used = _warm_pyramid(core, monkeypatch, pair, _FixedNet(bad), (1,))
    assert not used
    assert w.warm_objectives[0] > w.warm_objectives[1]
    c = fr.FastRegistration(voxels=pair.fixed, nodes=pair.moving,
                            num_voxel_intesity_levels=32, num_node_intensity_levels=32,
                            gamma=1e-6, convergence_th=2.5e-5, topology_correction=True,
                            increase_gamma_at_DS_resolution=True)
    c.run([8, 4, 2, 1], 10)
    assert [lv[1] for lv in w.level_log] == [lv[1] for lv in c.level_log]
    np.testing.assert_array_equal(w.deformedVoxelPositions, c.deformedVoxelPositions)
