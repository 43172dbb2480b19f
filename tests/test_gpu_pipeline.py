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


@pytest.mark.parametrize("shape_in,shape_out,comps", [
    ((160, 192, 160), (20, 24, 20), 3),   # first level of the cold pyramid
    ((17, 9, 13), (9, 18, 5), 3),         # mixed shrink / grow, odd sizes
    ((96, 80, 1), (12, 10, 1), 2)])       # 2D
def test_resize_index_grid(core, shape_in, shape_out, comps):
    """The index grid resampled axis by axis (each component's 1-D ramp
    resized with the same kernels, then broadcast) equals the uploaded
    index grid resampled as a volume up to FP64 rounding (the Gaussian and
    the zoom along the other axes scale a constant by a sum of weights that
    is 1 to about 1e-16), equals it after the float32 rounding the
    registration stores, and matches the host restatement of
    skimage.transform.resize(order=1)."""
    from binder_amd import resample
    print(f"resize index grid {shape_in} -> {shape_out} x{comps}", flush=True)
    grid = np.stack(np.indices(shape_in, dtype=np.float64), -1)[..., :comps]
    scale = shape_out[0] / shape_in[0]
    got = core.resize_index_grid(shape_in, comps, shape_out, scale)
    assert got.shape == shape_out + (comps,)
    vol = core.resize_field(grid, shape_out, scale)
    assert np.max(np.abs(got - vol)) <= 1e-13 * np.abs(vol).max()
    assert np.array_equal(got.astype(np.float32), vol.astype(np.float32))
    exp = resample.resize(grid, shape_out, order=1) * scale
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


@pytest.mark.parametrize("shape_in,shape_out,scale", [
    ((20, 24, 20), (40, 48, 40), 2.0),
    ((80, 96, 80), (160, 192, 160), 2.0),
    ((17, 9, 13), (9, 18, 5), 0.5),
    ((30, 22, 1), (15, 44, 1), 1.0)])
def test_resampled_registration_matches_host_path(core, shape_in, shape_out, scale):
    """Registration.resampled_from (resize on the GPU, straight into the new
    level's registration) holds exactly the field of the host path:
    Registration(resize_field(src.get_final_locations_as_numpy_array()))."""
    from binder_amd import synth
    print(f"resampled registration {shape_in} -> {shape_out}", flush=True)
    N = 2 if shape_in[2] == 1 else 3
    rng = np.random.default_rng(sum(shape_in) + sum(shape_out))
    vin = np.stack(np.indices(shape_in, dtype=np.float64), -1)[..., :N]
    f = vin + rng.standard_normal(vin.shape) * 1.5
    src = core.Registration(f.ravel(), 1, vin.ravel(), None, *shape_in, *shape_in, 3,
                            float(np.prod(shape_in)))
    vout = np.stack(np.indices(shape_out, dtype=np.float64), -1)[..., :N]
    a = core.Registration.resampled_from(src, scale, vout.ravel(), *shape_out, *shape_out, 3,
                                         float(np.prod(shape_out)))
    host = core.resize_field(np.asarray(src.get_final_locations_as_numpy_array(), np.float64),
                             shape_out, scale)
    b = core.Registration(host.ravel(), 1, vout.ravel(), None, *shape_out, *shape_out, 3,
                          float(np.prod(shape_out)))
    ga = a.get_final_locations_as_numpy_array()
    gb = b.get_final_locations_as_numpy_array()
    assert ga.shape == gb.shape == tuple(shape_out) + (N,)
    assert np.array_equal(ga, gb)


def test_pyramid_device_chain_matches_host_chain(core, monkeypatch):
    """The pyramid with the field kept on the GPU between levels (device
    resize and topology correction) against the same pyramid with a host
    round trip at every level: same iteration counts; fields equal up to the
    float32 rounding of the topology-corrected field that the device chain
    keeps (the host chain carries it in float64 into the next resize)."""
    from binder_amd import synth
    import binder_amd.FastRegistration as fr
    print("pyramid 48x40x44 device chain vs host chain", flush=True)
    pair = synth.make_pair((48, 40, 44), 5, 2.0, seed=3, semi=(3.0, 12.0))
    g = _pyramid(core, monkeypatch, pair)
    monkeypatch.delattr(core.Registration, "resampled_from")
    h = _pyramid(core, monkeypatch, pair)
    gi = [lv[1] for lv in g.level_log]
    hi = [lv[1] for lv in h.level_log]
    diff = np.asarray(g.deformedVoxelPositions, np.float64) - \
        np.asarray(h.deformedVoxelPositions, np.float64)
    rms = float(np.sqrt(np.mean(np.sum(diff ** 2, -1))))
    print(f"iterations {gi} / {hi}; RMS diff {rms:.2e}, max {np.abs(diff).max():.2e}",
          flush=True)
    assert gi == hi
    assert rms < 0.01


def test_resampled_registration_index_grid(core):
    """voxel_pos=None (the index grid, generated on the GPU) gives the same
    EM trajectory as passing the index grid from the host."""
    from binder_amd import synth
    print("resampled registration with a generated index grid", flush=True)
    pair, prob = synth.config_problem("c2", seed=4, scale=0.2)
    Nx, Ny, Nz = prob.shape
    src = synth.build_registration(core, prob)
    a = core.Registration.resampled_from(src, 1.0, prob.voxel_pos.ravel(), Nx, Ny, Nz,
                                         Nx, Ny, Nz, 3, prob.nnz)
    b = core.Registration.resampled_from(src, 1.0, None, Nx, Ny, Nz, Nx, Ny, Nz, 3, prob.nnz)
    for r in (a, b):
        r.setNonLinearTransformation(*prob.gammas, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, prob.D.ravel())
        r.setMILikelihood(prob.alpha, prob.theta.ravel(), prob.K, prob.L, 1,
                          prob.binned_nodes.ravel(), prob.binned_voxels.ravel())
        r.EM(3, -1e30, 1, 0)
    assert np.array_equal(a.get_final_locations_as_numpy_array(),
                          b.get_final_locations_as_numpy_array())
    assert np.array_equal(src.get_final_locations_as_numpy_array(),
                          prob.final_locations.astype(np.float32))


@pytest.mark.parametrize("shape", [(24, 20, 18), (40, 36, 1)])
def test_device_prior_spectrum_matches_uploaded(core, shape):
    """setNonLinearTransformationDCTPrior (D made on the GPU) gives exactly
    the EM trajectory of setNonLinearTransformation with the host D."""
    from binder_amd import synth
    print(f"device-made D, {shape}", flush=True)
    nd = 2 if shape[2] == 1 else 3
    pair = synth.make_pair(shape[:nd], 4, 1.5, seed=2, semi=(3.0, 8.0))
    prob = synth.problem_from_pair(pair, gamma=1e-4)
    vs = (1.0, 1.0, 1.0)
    D = core.dct_prior_eigenvalues(prob.shape[:nd], 2, nd, voxel_size=vs)
    fields = []
    for device in (False, True):
        r = synth.build_registration(core, prob)
        if device:
            r.setNonLinearTransformationDCTPrior(*prob.gammas, *vs, *vs, 2)
        else:
            r.setNonLinearTransformation(*prob.gammas, *vs, *vs, D.ravel())
        r.EM(6, -1e30, 1, 0)
        fields.append((r.get_final_locations_as_numpy_array(), r.get_objective_trace()))
    np.testing.assert_array_equal(fields[0][0], fields[1][0])
    np.testing.assert_array_equal(fields[0][1], fields[1][1])
