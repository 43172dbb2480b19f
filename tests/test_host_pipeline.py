"""Host-side pipeline pieces (no GPU): NIfTI I/O, resampling, masks and the
FastRegistration pyramid driven by the FP64 oracle backend."""
import os

import numpy as np
import pytest

from binder_amd import nifti, resample, utils

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_nifti_reads_reference_test_images():
    a = nifti.load(os.path.join(GOLDEN, "T1w_1_8mm.nii.gz"))
    b = nifti.load(os.path.join(GOLDEN, "T1w_2_8mm.nii.gz"))
    assert a.shape == b.shape == (22, 24, 24)
    # sform and qform of these files describe the same 8 mm rotated grid
    assert np.allclose(np.sqrt((a.affine[:3, :3] ** 2).sum(0)), 8.0, atol=1e-3)
    assert a.get_fdata().dtype == np.float64 and a.get_fdata().max() > 100


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.uint8])
def test_nifti_roundtrip(tmp_path, dtype):
    rng = np.random.default_rng(3)
    data = (rng.standard_normal((5, 6, 7, 3)) * 50).astype(dtype)
    A = np.array([[1.5, 0.1, 0, 3], [0, 2, 0.1, -1], [0.2, 0, 1, 2], [0, 0, 0, 1.0]])
    p = str(tmp_path / "x.nii.gz")
    nifti.save(nifti.NiftiImage(data, A), p)
    back = nifti.load(p)
    assert back.dataobj.dtype == dtype and np.array_equal(back.dataobj, data)
    assert np.allclose(back.affine, A)


def test_resize_constant_identity_and_clip():
    rng = np.random.default_rng(4)
    f = np.full((12, 10, 8, 3), 2.5)
    assert np.allclose(resample.resize(f, (6, 5, 4)), 2.5)
    assert np.allclose(resample.resize(f, (24, 20, 16)), 2.5)
    g = rng.standard_normal((9, 7, 5, 2))
    assert np.array_equal(resample.resize(g, (9, 7, 5)), g)
    up = resample.resize(g, (18, 14, 10))
    assert up.shape == (18, 14, 10, 2) and up.min() >= g.min() and up.max() <= g.max()
    # linear interpolation on the pixel-extent grid: doubling a linear ramp
    # gives the ramp sampled at (i + 0.5) / 2 - 0.5 away from the mirrored edges
    ramp = np.arange(10, dtype=np.float64)[:, None, None] * np.ones((1, 3, 3))
    r2 = resample.resize(ramp, (20, 3, 3))
    i = np.arange(2, 18)
    assert np.allclose(r2[i, 1, 1], (i + 0.5) / 2 - 0.5)


def test_warp_identity_and_nearest():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 4, (7, 6, 5)).astype(np.float64)
    pos = np.indices(img.shape, dtype=np.float64)
    assert np.array_equal(resample.warp(img, pos, order=0), img)
    assert np.allclose(resample.warp(img, pos + 0.3, order=0), img)  # rounds back
    shifted = resample.warp(img, pos + np.array([1, 0, 0])[:, None, None, None], order=0)
    assert np.array_equal(shifted[:-1], img[1:]) and np.all(shifted[-1] == 0)


def test_mask_and_hist_norm():
    img = np.zeros((12, 12, 12))
    img[3:9, 3:9, 3:9] = 5.0
    img[5, 5, 5] = 0.0  # a hole
    m = utils.maskImage(img, th_value=1.0)
    assert m.dtype == np.int32 and m[5, 5, 5] == 1 and m.sum() == 216
    rng = np.random.default_rng(6)
    src = rng.random(1000)
    tmpl = rng.random(1000) * 10 + 5
    out = utils.hist_norm(src, tmpl)
    assert np.argsort(out, kind="stable").tolist() == np.argsort(src, kind="stable").tolist()
    assert out.min() >= tmpl.min() - 1e-12 and out.max() <= tmpl.max() + 1e-12


FIXED_2D = np.array([[0, 0, 0, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
                     [0, 0, 0, 1, 1, 2, 3, 1, 0, 0], [0, 0, 1, 1, 1, 2, 3, 1, 0, 0],
                     [0, 0, 1, 1, 1, 2, 3, 1, 0, 0], [0, 0, 1, 1, 1, 2, 3, 1, 0, 0],
                     [0, 0, 1, 1, 1, 2, 3, 1, 0, 0], [0, 0, 1, 1, 1, 1, 1, 1, 0, 0],
                     [0, 0, 0, 1, 1, 1, 1, 1, 0, 0], [0, 0, 0, 0, 0, 1, 1, 1, 0, 0],
                     [0, 0, 0, 0, 0, 0, 0, 0, 0, 0]])
MOVING_2D = np.array([[0, 0, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0, 0, 0],
                      [0, 0, 0, 0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0, 0, 0],
                      [0, 0, 0, 1, 1, 2, 3, 1, 0], [0, 0, 1, 1, 1, 2, 3, 1, 0],
                      [0, 0, 1, 1, 1, 2, 3, 1, 0], [0, 0, 1, 1, 1, 2, 3, 1, 0],
                      [0, 0, 1, 1, 1, 2, 3, 1, 0], [0, 0, 1, 1, 1, 1, 1, 1, 0],
                      [0, 0, 0, 1, 1, 1, 1, 1, 0], [0, 0, 0, 0, 0, 1, 1, 1, 0],
                      [0, 0, 0, 0, 0, 0, 0, 0, 0]])


def toy_2d_registration(backend_module, monkeypatch):
    """BINDER/tests/cpu/test_2D_MI_registration.py:7-80 on a given backend."""
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", backend_module)
    reg = fr.FastRegistration(voxels=FIXED_2D, nodes=MOVING_2D, num_voxel_intesity_levels=4,
                              num_node_intensity_levels=4, gamma=100, threads=1, metric='MI',
                              seed=12345, alpha=1.1, bin_robust=False,
                              topology_correction=False)
    reg.run(resolutions=[1], max_number_of_iterations=100)
    pos = np.moveaxis(reg.deformedVoxelPositions, 3, 0)[0:2, :, :, 0]
    warped_em = resample.warp(MOVING_2D, pos, order=0)
    mean, _ = reg.sample_param(0, 100)
    pos = np.moveaxis(np.asarray(mean), 3, 0)[0:2, :, :, 0]
    warped_mcmc = resample.warp(MOVING_2D, pos, order=0)
    return warped_em, warped_mcmc


def test_toy_2d_registration_oracle(orc, monkeypatch):
    """The oracle recovers the toy deformation exactly, EM and MCMC mean."""
    em, mcmc = toy_2d_registration(orc, monkeypatch)
    assert np.allclose(em, FIXED_2D)
    assert np.allclose(mcmc, FIXED_2D)


def test_fastregistration_rejects_out_of_scope_options(orc, monkeypatch):
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", orc)
    for kw in ({"metric": "NCC"}, {"feature_moving": True}, {"metric": "SSD", "feature_fixed": True},
               {"smooth": True}):
        with pytest.raises(NotImplementedError):
            fr.FastRegistration(voxels=FIXED_2D, nodes=MOVING_2D, topology_correction=False, **kw)
    with pytest.raises(ValueError):
        fr.FastRegistration(voxels=FIXED_2D, nodes=MOVING_2D, topology_correction=False,
                            spline_order=4)


def test_pyramid_bookkeeping(orc, monkeypatch):
    """move_resolution: shapes ceil(N/r), positions / r, gamma scaled by the
    voxel volume, voxel count and squared node size (with and without the
    increase at coarse levels), field rescaled by current/new resolution."""
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", orc)
    rng = np.random.default_rng(8)
    fixed = rng.random((17, 14, 11))
    moving = rng.random((17, 14, 11))
    for inc in (False, True):
        reg = fr.FastRegistration(voxels=fixed, nodes=moving, gamma=1e-6,
                                  topology_correction=False, num_voxel_intesity_levels=8,
                                  num_node_intensity_levels=8,
                                  increase_gamma_at_DS_resolution=inc)
        reg.move_resolution(4)
        assert reg.voxels_shape == (5, 4, 3)
        assert reg.deformedVoxelPositions.shape == (5, 4, 3, 3)
        assert np.allclose(reg.voxels_pos, reg.full_voxels_pos[::4, ::4, ::4] / 4)
        count = np.prod((17, 14, 11)) if inc else np.prod((5, 4, 3))
        assert np.isclose(reg.gamma_x, 1e-6 * 4 ** 3 * count * 4 ** 2)
        assert reg.D.shape == (5, 4, 3) and reg.D[0, 0, 0] == 0.0
        # identity at full resolution maps to (about) the identity at 1/4
        assert np.abs(reg.deformedVoxelPositions - reg.voxels_pos)[1:-1, 1:-1, 1:-1].max() < 0.3


def t1_registration(backend_module, monkeypatch, tmp_path):
    """BINDER/tests/cpu/test_3D_MI_registration.py:29-96 on a given backend:
    3 EM iterations on the 8 mm T1 pair, then 20 sampler sweeps; returns the
    SSDs of the moving image, the EM warp and the sample-mean warp."""
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", backend_module)
    a = nifti.load(os.path.join(GOLDEN, "T1w_1_8mm.nii.gz"))
    b = nifti.load(os.path.join(GOLDEN, "T1w_2_8mm.nii.gz"))
    fixed, moving = a.get_fdata(), b.get_fdata()
    reg = fr.FastRegistration(voxels=fixed, nodes=moving, voxels_matrix=a.affine,
                              nodes_matrix=b.affine, num_voxel_intesity_levels=32,
                              num_node_intensity_levels=32, gamma=1e-4, metric='MI', threads=1,
                              spline_order=3, topology_correction=False)
    reg.run(resolutions=[1], max_number_of_iterations=3)
    utils.save_warped_image(deformed_position=reg.deformedVoxelPositions, movingImage=moving,
                            outputDir=str(tmp_path), affine=a.affine)
    warped = nifti.load(str(tmp_path / "warped.nii.gz")).get_fdata()
    mean, _ = reg.sample_param(burnin=0, samples=20, sample_gamma=False)
    utils.save_warped_image(deformed_position=np.asarray(mean), movingImage=moving,
                            outputDir=str(tmp_path), affine=a.affine,
                            outputName="mean_warped.nii.gz")
    warped_mc = nifti.load(str(tmp_path / "mean_warped.nii.gz")).get_fdata()
    return (np.sum((moving - fixed) ** 2), np.sum((warped - fixed) ** 2),
            np.sum((warped_mc - fixed) ** 2))


def test_t1_registration_oracle(orc, monkeypatch, tmp_path):
    start, em, mc = t1_registration(orc, monkeypatch, tmp_path)
    assert em < start and mc < start and mc < em


def run_cli(tmp_path, extra):
    from binder_amd.cli import run_binder
    out = tmp_path / "out"
    args = ["-f", os.path.join(GOLDEN, "T1w_1_8mm.nii.gz"),
            "-m", os.path.join(GOLDEN, "T1w_2_8mm.nii.gz"), "-o", str(out),
            "--resolutions", "2", "1", "--iterations", "4", "--bins-fixed", "32",
            "--bins-moving", "32", "--sampler", "--number-of-burnin", "2",
            "--number-of-samples", "4", "--save-warped-image", "--save-binned-images",
            "--save-masks", "--save-likelihood-parameters"] + extra
    assert run_binder.main(args) == 0
    shape = (22, 24, 24)
    for name, shp in [("warp.nii.gz", shape + (3,)), ("deformation_field.nii.gz", shape + (3,)),
                      ("warped.nii.gz", shape), ("binned_warped.nii.gz", shape),
                      ("binned_fixed.nii.gz", shape), ("mean_warp_sample.nii.gz", shape + (3,)),
                      ("mean_deformation_field_sample.nii.gz", shape + (3,)),
                      ("covariance_field_sample.nii.gz", shape + (3, 3)),
                      ("warped_sample.nii.gz", shape), ("mask_fixed.nii.gz", shape),
                      ("mask_moving.nii.gz", shape)]:
        img = nifti.load(str(out / name))
        assert img.shape == shp, name
        assert np.all(np.isfinite(img.get_fdata())), name
    assert (out / "likelihood_parameters.npz").exists()
    return out


def test_cli_oracle(orc, monkeypatch, tmp_path):
    """run_BINDER's flags and outputs (BINDER/cli/run_BINDER.py), oracle backend."""
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", orc)
    run_cli(tmp_path, ["--no-topology-correction"])


@pytest.mark.parametrize("kw", [{"affine": True}, {"rigid": True}], ids=["affine", "rigid"])
def test_toy_2d_linear_registration_oracle(orc, monkeypatch, kw):
    """The toy pair differs by a translation of two rows: the affine and the
    rigid fit (MI likelihood) recover it and the warped moving image equals
    the fixed one."""
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", orc)
    reg = fr.FastRegistration(voxels=FIXED_2D, nodes=MOVING_2D, num_voxel_intesity_levels=4,
                              num_node_intensity_levels=4, gamma=100, threads=1,
                              seed=12345, alpha=1.1, bin_robust=False,
                              topology_correction=False, **kw)
    reg.run(resolutions=[1], max_number_of_iterations=100)
    pos = np.moveaxis(reg.deformedVoxelPositions, 3, 0)[0:2, :, :, 0]
    assert np.allclose(resample.warp(MOVING_2D, pos, order=0), FIXED_2D)
    par = np.asarray(reg.registratorObj.get_linear_parameters())
    assert abs(par[-2] - 2.0) < 0.2 and abs(par[-1]) < 0.25


def test_toy_2d_ssd_hand_over(orc, monkeypatch):
    """metric='SSD': sigma^2 starts at the mean squared bin difference over
    the common extent (FastRegistration.py:241-246) and a level hands
    get_likelihood_parameters() = [scale, shift] on as (shift, scale), as
    the reference's unpacking at FastRegistration.py:521 does."""
    import binder_amd.FastRegistration as fr
    monkeypatch.setattr(fr, "Registration", orc)
    reg = fr.FastRegistration(voxels=FIXED_2D, nodes=MOVING_2D, num_voxel_intesity_levels=4,
                              num_node_intensity_levels=4, gamma=100, threads=1, metric="SSD",
                              alpha_0=1, beta_0=1, bin_robust=False, topology_correction=False)
    bf = utils.bin_signal(FIXED_2D[:, :, None] * 1.0, 4)
    bm = utils.bin_signal(MOVING_2D[:, :, None] * 1.0, 4)
    n = np.minimum(bf.shape, bm.shape)
    want = np.sum((bm[:n[0], :n[1], :n[2]] - bf[:n[0], :n[1], :n[2]]) ** 2) / np.prod(n)
    assert np.isclose(reg.sigma_sq, want)
    reg.run(resolutions=[1], max_number_of_iterations=20)
    scale, shift = reg.registratorObj.get_likelihood_parameters()
    assert (reg.shift, reg.scale) == (scale, shift)
