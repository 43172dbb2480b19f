"""The reference's registration tests and a full pyramid on the HIP backend
(BINDER/tests/cpu/test_2D_MI_registration.py, test_3D_MI_registration.py,
BINDER/FastRegistration.py:443-650)."""
import numpy as np
import pytest

from test_host_pipeline import FIXED_2D, t1_registration, toy_2d_registration

pytestmark = pytest.mark.gpu


def test_toy_2d_registration_gpu(core, monkeypatch):
    """Exact recovery of the toy deformation (EM and sample mean)."""
    print("toy 2D registration (HIP)", flush=True)
    em, mcmc = toy_2d_registration(core, monkeypatch)
    assert np.allclose(em, FIXED_2D)
    assert np.allclose(mcmc, FIXED_2D)


def test_t1_registration_gpu(core, monkeypatch, tmp_path):
    """EM and sampling both lower the SSD of the 8 mm T1 pair; sampling
    lowers it below the 3-iteration EM result."""
    print("T1 8mm registration (HIP)", flush=True)
    start, em, mc = t1_registration(core, monkeypatch, tmp_path)
    assert em < start and mc < start and mc < em


def _pyramid(backend_module, monkeypatch, pair):
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", backend_module)
    reg = fr.FastRegistration(voxels=pair.fixed, nodes=pair.moving,
                              num_voxel_intesity_levels=32, num_node_intensity_levels=32,
                              gamma=1e-6, convergence_th=2.5e-5, topology_correction=True,
                              increase_gamma_at_DS_resolution=True)
    reg.run(resolutions=[4, 2, 1], max_number_of_iterations=10)
    return reg


def test_pyramid_matches_oracle(core, orc, monkeypatch):
    """Three-level pyramid with topology correction on a 48x40x44 synthetic
    pair: same EM iteration counts per level as the FP64 oracle and a final
    field within 0.025 voxel RMS of it, a quarter of the 0.1-voxel RMS
    time-to-tolerance criterion.  (FP32 rounding moves the field by about
    1e-5 per iteration; the discrete fold decisions of the topology
    correction turn that into sub-voxel changes at a few borderline cells.)"""
    from binder_amd import synth
    print("pyramid 48x40x44 (HIP vs oracle)", flush=True)
    pair = synth.make_pair((48, 40, 44), 5, 2.0, seed=3, semi=(3.0, 12.0))
    g = _pyramid(core, monkeypatch, pair)
    o = _pyramid(orc, monkeypatch, pair)
    gi = [lv[1] for lv in g.level_log]
    oi = [lv[1] for lv in o.level_log]
    diff = np.asarray(g.deformedVoxelPositions, np.float64) - o.deformedVoxelPositions
    rms = float(np.sqrt(np.mean(np.sum(diff ** 2, -1))))
    print(f"iterations HIP {gi} oracle {oi}; field RMS diff {rms:.2e}, max {np.abs(diff).max():.2e}",
          flush=True)
    assert gi == oi
    assert rms < 0.025


def test_cli_gpu(core, tmp_path):
    """run_BINDER end to end on the HIP backend, with topology correction and
    the Jacobian output."""
    from test_host_pipeline import run_cli
    from binder_amd import nifti
    print("run_binder CLI (HIP)", flush=True)
    out = run_cli(tmp_path, ["--save-jacobian-determinant"])
    jac = nifti.load(str(out / "jacobian_determinant.nii.gz")).get_fdata()
    assert jac.shape == (22, 24, 24) and np.all(np.isfinite(jac))


@pytest.mark.parametrize("shape_in,shape_out,comps", [
    ((160, 192, 160), (20, 24, 20), 3),   # first pyramid level from full size
    ((20, 24, 20), (40, 48, 40), 3),      # level up-sampling
    ((80, 96, 80), (160, 192, 160), 3),
    ((17, 9, 13), (9, 18, 5), 1),         # mixed shrink / grow, odd sizes
    ((30, 22, 1), (15, 44, 1), 2)])       # 2D field
def test_resize_field_matches_host(core, shape_in, shape_out, comps):
    """GPU pyramid resampling equals the host restatement of
    skimage.transform.resize(order=1) (binder_amd.resample.resize)."""
    from binder_amd import resample
    print(f"resize {shape_in} -> {shape_out} x{comps}", flush=True)
    rng = np.random.default_rng(sum(shape_in))
    f = rng.standard_normal(shape_in + (comps,)) * 10
    f += np.stack(np.indices(shape_in, dtype=np.float64), -1)[..., :comps]
    got = core.resize_field(f, shape_out, 0.5)
    exp = resample.resize(f, shape_out, order=1) * 0.5
    assert got.shape == exp.shape
    assert np.max(np.abs(got - exp)) <= 1e-10 * np.abs(exp).max()


def test_pyramid_2d_matches_oracle(core, orc, monkeypatch):
    """2D pyramid (96x80 synthetic pair, 6 ellipses, 3 px warp) with 2D
    topology correction: same per-level iteration counts as the FP64 oracle
    and a final field within 0.025 pixel RMS (as in 3D)."""
    from binder_amd import synth
    print("pyramid 2D 96x80 (HIP vs oracle)", flush=True)
    pair = synth.make_pair((96, 80), 6, 3.0, seed=5, semi=(5.0, 30.0))
    g = _pyramid(core, monkeypatch, pair)
    o = _pyramid(orc, monkeypatch, pair)
    gi = [lv[1] for lv in g.level_log]
    oi = [lv[1] for lv in o.level_log]
    diff = np.asarray(g.deformedVoxelPositions, np.float64) - o.deformedVoxelPositions
    rms = float(np.sqrt(np.mean(np.sum(diff ** 2, -1))))
    print(f"iterations HIP {gi} oracle {oi}; field RMS diff {rms:.2e}", flush=True)
    assert gi == oi
    assert rms < 0.025
