"""HIP Jacobian determinant / fold detection (bit-exact) and topology
correction against the FP64 oracle (BINDER/src/cpu/utils.cpp:10-345,
BINDER/src/cpu/topology_correction_3D.cpp:17-607)."""
import numpy as np
import pytest

from binder_amd import synth

pytestmark = pytest.mark.gpu


def _folded_field(shape, rms, seed):
    rng = np.random.default_rng(seed)
    pos = np.stack(np.indices(shape, dtype=np.float64), axis=-1)
    return pos + synth.dct_displacement(shape, rms, rng, ncoef=4)


@pytest.mark.parametrize("shape,vs", [((7, 6, 5), (1.0, 1.0, 1.0)),
                                      ((33, 29, 31), (1.1, 0.9, 1.3)),
                                      ((64, 48, 40), (1.0, 1.0, 1.0)),
                                      ((2, 3, 2), (2.0, 1.0, 0.5))])
def test_jacobian_3d_bitexact(core, orc, shape, vs):
    print(f"jacobian 3D shape={shape} vs={vs}", flush=True)
    f = _folded_field(shape, 2.5, seed=shape[0])
    jg = core.compute_jacobian_determinant_3D(f.ravel(), *shape, *vs)
    jo = orc.compute_jacobian_determinant_3D(f.ravel(), *shape, *vs)
    assert np.array_equal(jg, jo)
    mask, cnt = core.fold_mask_3D(f.ravel(), *shape, *vs, e1=0.0)
    assert cnt == int((jo < 0).sum())
    assert np.array_equal(mask, jo < 0)


@pytest.mark.parametrize("shape", [(9, 7), (50, 41), (3, 2)])
def test_jacobian_2d_bitexact(core, orc, shape):
    print(f"jacobian 2D shape={shape}", flush=True)
    rng = np.random.default_rng(shape[0])
    pos = np.stack(np.indices(shape, dtype=np.float64), axis=-1)
    f = pos + rng.normal(0, 0.6, pos.shape)
    jg = core.compute_jacobian_determinant_2D(f.ravel(), *shape, 1.2, 0.8)
    jo = orc.compute_jacobian_determinant_2D(f.ravel(), *shape, 1.2, 0.8)
    assert np.array_equal(jg, jo)


def test_jacobian_affine_closed_form(core):
    """An affine map x -> A x + b has J = det(A) everywhere."""
    print("jacobian affine", flush=True)
    A = np.array([[1.1, 0.2, -0.1], [0.05, 0.9, 0.3], [0.0, -0.2, 1.2]])
    shape = (8, 9, 10)
    pos = np.stack(np.indices(shape, dtype=np.float64), axis=-1)
    f = pos @ A.T + np.array([3.0, -2.0, 0.5])
    J = core.compute_jacobian_determinant_3D(f.ravel(), *shape, 1, 1, 1)
    assert np.allclose(J, np.linalg.det(A), rtol=1e-12)


def test_topology_identity_untouched(core):
    print("topology identity", flush=True)
    shape = (10, 11, 12)
    pos = np.stack(np.indices(shape, dtype=np.float64), axis=-1)
    t = core.topology_correction_3D()
    out = t.correct_topology(pos.ravel(), 0.01, 100, 10, 10, 1, np.ones(3), *shape)
    assert np.array_equal(out, pos)
    assert t.last_stats[0] == 0


@pytest.mark.parametrize("shape,vs,rms", [((16, 14, 12), (1.0, 1.0, 1.0), 3.0),
                                          ((24, 20, 22), (1.2, 1.0, 0.8), 3.5)])
def test_topology_matches_oracle(core, orc, shape, vs, rms):
    print(f"topology shape={shape} vs={vs} rms={rms}", flush=True)
    f = _folded_field(shape, rms, seed=42 + shape[0])
    vs = np.array(vs)
    J0 = orc.compute_jacobian_determinant_3D(f.ravel(), *shape, *vs)
    assert (J0 < 0.01).sum() > 0, "test field must contain folds"
    tg = core.topology_correction_3D()
    to = orc.topology_correction_3D()
    og = tg.correct_topology(f.ravel(), 0.01, 100, 10, 10, 1, vs, *shape)
    oo = to.correct_topology(f.ravel(), 0.01, 100, 10, 10, 1, vs, *shape)
    # same outer/inner iteration counts and final fold counts
    assert np.array_equal(tg.last_stats[[0, 1, 4, 5]], to.last_stats[[0, 1, 4, 5]])
    # corrected field: the Poisson solves differ only by rounding (Hartley
    # contraction vs FFT)
    assert np.max(np.abs(og - oo)) <= 1e-6
    # mean position preserved by the integration step
    assert np.allclose(og.reshape(-1, 3).mean(0), f.reshape(-1, 3).mean(0), atol=1e-9)
    Jg = core.compute_jacobian_determinant_3D(og.ravel() * np.tile(vs, og.size // 3),
                                              *shape, *vs)
    Jo = orc.compute_jacobian_determinant_3D(oo.ravel() * np.tile(vs, oo.size // 3),
                                             *shape, *vs)
    assert (Jg < 0.01).sum() == (Jo < 0.01).sum()


def test_topology2d_identity_untouched(core):
    print("topology 2D identity", flush=True)
    shape = (20, 17)
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    t = core.topology_correction_2D()
    out = t.correct_topology(pos.ravel(), 0.01, 100, 10, 10, 1, np.ones(2), *shape)
    assert np.array_equal(out, pos) and t.last_stats[0] == 0


@pytest.mark.parametrize("shape,vs,rms,ncoef", [((24, 20), (1.0, 1.0), 2.5, 4),
                                                ((40, 33), (1.2, 0.8), 7.0, 4),
                                                ((192, 160), (1.0, 1.0), 6.0, 12)])
def test_topology2d_matches_oracle(core, orc, shape, vs, rms, ncoef):
    """Same iteration counts as the FP64 oracle and the same field to 1e-6
    (topology_correction_2D.cpp:17-430); every corner ends in [e1, e2]."""
    from binder_amd import synth
    print(f"topology 2D {shape} vs {vs}", flush=True)
    rng = np.random.default_rng(shape[0])
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    f = pos + synth.dct_displacement(shape, rms, rng, ncoef=ncoef)
    vs = np.array(vs)
    g = core.topology_correction_2D()
    o = orc.topology_correction_2D()
    a = g.correct_topology(f.ravel(), 0.01, 100, 10, 10, 1, vs, *shape)
    b = o.correct_topology(f.ravel(), 0.01, 100, 10, 10, 1, vs, *shape)
    assert g.last_stats[0] == o.last_stats[0] and g.last_stats[1] == o.last_stats[1]
    assert g.last_stats[0] > 0
    assert np.max(np.abs(a - b)) <= 1e-6
    _, cnt = orc.corner_jacobians_2D((a * vs).ravel(), *shape, vs, 0.01, 100)
    if g.last_stats[4] == 0:
        assert cnt[0] == 0
