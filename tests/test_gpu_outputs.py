"""The files run_BINDER writes (BINDER/cli/run_BINDER.py:252-307), from the
HIP backend against the FP64 oracle running the same command on the 8 mm T1
pair: the registered field and warp, the Jacobian output, the warped and
binned images (BINDER/utils.py:222-227), masks, likelihood parameters, and
the sampler's mean-warp and covariance files.

* Headers are identical (shape, affine, voxel sizes, datatype; field files
  are float32 from this backend as from the reference's GPU bindings, and
  float64 from the oracle as from its CPU bindings).
* Host-side outputs of the same inputs (binned fixed image, masks) are
  identical.
* Each file is checked against its own definition from the written field:
  the Jacobian file is the oracle's Jacobian of the written deformation
  field (bit for bit), the warped images are the image warped by the
  written field.
* Backend against oracle: the EM field within 2e-3 voxel (max) after 4 + 2
  iterations (levels 2 and 1) with topology correction, theta within 1e-4, Jacobians and
  warped images within what that field difference moves them.
* The sampler's files (Philox streams here, mt19937 in the oracle: equal in
  distribution only): mean fields within 4 Monte Carlo standard errors of
  each other (the error from the chains' own covariance, 200 samples, an
  allowance of 10 for autocorrelation), mean covariance traces within 25 %.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SHAPE = (22, 24, 24)


def _cli(tmp_path, name):
    from binder_amd.cli import run_binder
    out = tmp_path / name
    args = ["-f", os.path.join(GOLDEN, "T1w_1_8mm.nii.gz"),
            "-m", os.path.join(GOLDEN, "T1w_2_8mm.nii.gz"), "-o", str(out),
            "--resolutions", "2", "1", "--iterations", "2", "--bins-fixed", "32",
            "--bins-moving", "32", "--sampler", "--number-of-burnin", "20",
            "--number-of-samples", "200", "--save-warped-image", "--save-binned-images",
            "--save-masks", "--save-likelihood-parameters", "--save-jacobian-determinant"]
    assert run_binder.main(args) == 0
    return out


@pytest.fixture(scope="module")
def outputs(core, orc, tmp_path_factory):
    import binder_amd._core as core_mod
    import binder_amd.FastRegistration as fr
    tmp = tmp_path_factory.mktemp("cli")
    print("run_binder CLI, HIP backend", flush=True)
    hip = _cli(tmp, "hip")
    print("run_binder CLI, FP64 oracle backend", flush=True)
    mp = pytest.MonkeyPatch()
    try:
        mp.setattr(fr, "Registration", orc)
        mp.setattr(core_mod, "compute_jacobian_determinant_3D",
                   orc.compute_jacobian_determinant_3D)
        ora = _cli(tmp, "oracle")
    finally:
        mp.undo()
    return hip, ora


def _load(d, name):
    from binder_amd import nifti
    return nifti.load(str(d / name))


NAMES = ["warp.nii.gz", "deformation_field.nii.gz", "jacobian_determinant.nii.gz",
         "warped.nii.gz", "binned_warped.nii.gz", "binned_fixed.nii.gz",
         "mean_warp_sample.nii.gz", "mean_deformation_field_sample.nii.gz",
         "covariance_field_sample.nii.gz", "warped_sample.nii.gz", "mask_fixed.nii.gz",
         "mask_moving.nii.gz"]


# files written from a backend's float array: float32 from this backend (as
# from the reference's GPU bindings), float64 from the oracle (as from its
# CPU bindings)
FIELD_FILES = {"warp.nii.gz", "deformation_field.nii.gz", "mean_warp_sample.nii.gz",
               "mean_deformation_field_sample.nii.gz", "covariance_field_sample.nii.gz"}


@pytest.mark.parametrize("name", NAMES)
def test_headers_identical(outputs, name):
    hip, ora = outputs
    a, b = _load(hip, name), _load(ora, name)
    assert a.shape == b.shape and a.shape[:3] == SHAPE, name
    assert np.array_equal(a.affine, b.affine), name
    assert np.array_equal(a.header["pixdim"], b.header["pixdim"]), name
    if name in FIELD_FILES:
        assert {a.header["datatype"], b.header["datatype"]} <= {16, 64}, name
    else:
        assert a.header["datatype"] == b.header["datatype"], name
    assert np.all(np.isfinite(a.get_fdata())), name


def test_host_outputs_identical(outputs):
    hip, ora = outputs
    for name in ("binned_fixed.nii.gz", "mask_fixed.nii.gz", "mask_moving.nii.gz"):
        assert np.array_equal(_load(hip, name).get_fdata(), _load(ora, name).get_fdata()), name


def test_em_field_and_theta(outputs):
    hip, ora = outputs
    fh = _load(hip, "deformation_field.nii.gz").get_fdata()
    fo = _load(ora, "deformation_field.nii.gz").get_fdata()
    diff = np.abs(fh - fo)
    print(f"EM field: max diff {diff.max():.2e} voxel, RMS "
          f"{np.sqrt(np.mean(np.sum((fh - fo) ** 2, -1))):.2e}", flush=True)
    assert diff.max() < 2e-3
    # warp = field - positions in both files
    wh, wo = _load(hip, "warp.nii.gz").get_fdata(), _load(ora, "warp.nii.gz").get_fdata()
    assert np.allclose(fh - wh, fo - wo, atol=1e-5)
    th = np.load(str(hip / "likelihood_parameters.npz"), allow_pickle=False)
    to = np.load(str(ora / "likelihood_parameters.npz"), allow_pickle=False)
    assert sorted(th.files) == sorted(to.files)
    for k in th.files:
        if th[k].dtype.kind in "US":
            assert th[k] == to[k], k
        else:
            assert np.allclose(th[k], to[k], rtol=1e-4, atol=1e-6), k


def test_jacobian_file(outputs, orc):
    """The written Jacobian is the oracle's Jacobian of the written field
    (BINDER/cli/run_BINDER.py:260-270; voxel size 8 mm), bit for bit; the
    backends' Jacobians differ by what their fields' difference moves."""
    hip, ora = outputs
    img = _load(hip, "deformation_field.nii.gz")
    vs = np.sqrt(np.sum(img.affine[:3, :3] ** 2, 0))
    for d in (hip, ora):
        f = _load(d, "deformation_field.nii.gz").get_fdata()
        f = np.ascontiguousarray(f / vs[None, None, None, :])
        ref = orc.compute_jacobian_determinant_3D(f.ravel(), *SHAPE, *vs)
        jac = _load(d, "jacobian_determinant.nii.gz").get_fdata()
        assert np.array_equal(jac, np.asarray(ref).reshape(SHAPE))
    jh = _load(hip, "jacobian_determinant.nii.gz").get_fdata()
    jo = _load(ora, "jacobian_determinant.nii.gz").get_fdata()
    assert np.abs(jh - jo).max() < 0.02


def test_warped_images(outputs):
    """Warped images are the moving image warped by the written field
    (BINDER/utils.py:222-227, cubic as the command line's default); the
    backends' images agree to the intensity change their field difference
    causes."""
    from binder_amd import nifti, resample
    hip, ora = outputs
    moving = nifti.load(os.path.join(GOLDEN, "T1w_2_8mm.nii.gz")).get_fdata()
    for d in (hip, ora):
        f = _load(d, "deformation_field.nii.gz").get_fdata()
        w = _load(d, "warped.nii.gz").get_fdata()
        pos = np.moveaxis(f, -1, 0)
        ref = resample.warp(moving, pos, order=3)
        assert np.array_equal(w, ref)
    wh, wo = _load(hip, "warped.nii.gz").get_fdata(), _load(ora, "warped.nii.gz").get_fdata()
    g = np.sqrt(sum(x ** 2 for x in np.gradient(moving)))
    assert np.abs(wh - wo).max() <= 4e-3 * g.max() + 1e-6


def test_sampler_files(outputs):
    hip, ora = outputs
    mh = _load(hip, "mean_deformation_field_sample.nii.gz").get_fdata()
    mo = _load(ora, "mean_deformation_field_sample.nii.gz").get_fdata()
    ch = _load(hip, "covariance_field_sample.nii.gz").get_fdata()
    co = _load(ora, "covariance_field_sample.nii.gz").get_fdata()
    assert ch.shape == SHAPE + (3, 3)
    # covariance files are symmetric with a non-negative diagonal
    assert np.allclose(ch, np.swapaxes(ch, -1, -2)) and np.allclose(co, np.swapaxes(co, -1, -2))
    dh = np.trace(ch, axis1=-2, axis2=-1)
    do = np.trace(co, axis1=-2, axis2=-1)
    assert dh.min() >= 0 and do.min() >= 0
    print(f"mean covariance trace HIP {dh.mean():.4e}, oracle {do.mean():.4e}", flush=True)
    assert abs(dh.mean() - do.mean()) <= 0.25 * do.mean()
    se2 = (dh + do) / 200.0 * 10.0
    rms = np.sqrt(np.mean(np.sum((mh - mo) ** 2, -1)))
    print(f"sample means: RMS diff {rms:.3e} voxel, standard error "
          f"{np.sqrt(np.mean(se2)):.3e}", flush=True)
    assert rms <= 4 * np.sqrt(np.mean(se2))
    # mean warp = mean field - positions
    wh = _load(hip, "mean_warp_sample.nii.gz").get_fdata()
    wo = _load(ora, "mean_warp_sample.nii.gz").get_fdata()
    assert np.allclose(mh - wh, mo - wo, atol=1e-5)
