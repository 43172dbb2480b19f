"""run_BINDER on the gfx950 backend: the reference's command line
(BINDER/cli/run_BINDER.py:9-307) with the same flags, for the MI metric with
the non-linear transformation.

Fixed relative to the reference script: it reads options whose flags are
commented out there (scale/shift, histogram matching, alpha/beta zero,
feature and affine/rigid options) and defines `metric` only when no moving
mask is given; here those options take the constructor defaults and the
metric is always MI.  `--initial-transformation` accepts a 4x4 matrix as
text or .npy (the reference's surfa .mat/.lta reader is not available).

    python -m binder_amd.cli.run_binder -f fixed.nii.gz -m moving.nii.gz -o out/
"""
import argparse
import os
import time

import numpy as np


def parse_arguments(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    default_threads = int(os.environ.get("OMP_NUM_THREADS", 1))
    p.add_argument("-o", "--output", metavar="DIR", help="Output directory.", required=True)
    p.add_argument("-f", "--fixed", dest="fixed", help="Fixed image.", required=True)
    p.add_argument("-m", "--moving", dest="moving", help="Moving image.", required=True)
    p.add_argument("--mask-fixed", type=str, help="Mask to apply to the fixed image.")
    p.add_argument("--mask-moving", type=str, help="Mask to apply to the moving image.")
    p.add_argument("--automatic-mask", action="store_true", default=False,
                   help="Estimate mask for both images (see mask-below).")
    p.add_argument("--mask-below", type=float, default=0, help="Mask values below a threshold.")
    p.add_argument("--save-masks", action="store_true", default=False, help="Save masks.")
    p.add_argument("--iterations", type=int, default=1000,
                   help="Maximum number of iterations per resolution.")
    p.add_argument("--convergence-th", type=float, default=2.5e-5, help="Convergence threshold.")
    p.add_argument("--do-not-increase-gamma", action="store_true", default=False,
                   help="Do not increase gamma at each DS by DS^D.")
    p.add_argument("--save-warped-image", action="store_true", default=False)
    p.add_argument("--save-binned-images", action="store_true", default=False)
    p.add_argument("--save-jacobian-determinant", action="store_true", default=False)
    p.add_argument("--output-interpolation", type=int, default=3,
                   help="Interpolation order for warped image.")
    p.add_argument("--initial-transformation", type=str,
                   help="Initial transformation (4x4 matrix, text or .npy).")
    p.add_argument("--do-not-initialize-center-of-mass", action="store_true", default=False)
    p.add_argument("--MI", action="store_true", default=True,
                   help="Mutual information metric (the only one).")
    p.add_argument("--bins-fixed", type=int, default=64, help="Number of bins for the fixed image.")
    p.add_argument("--bins-moving", type=int, default=64,
                   help="Number of bins for the moving image.")
    p.add_argument("--do-not-bin-robust", action="store_true", default=False)
    p.add_argument("--alpha", type=float, default=2.0, help="Prior on theta (for each k).")
    p.add_argument("--no-topology-correction", action="store_true", default=False)
    p.add_argument("--e1", type=float, default=0.01, help="Minimum Jacobian determinant.")
    p.add_argument("--e2", type=float, default=100, help="Maximum Jacobian determinant.")
    p.add_argument("--max_outer_it", type=int, default=10)
    p.add_argument("--max_inner_it", type=int, default=10)
    p.add_argument("--resolutions", nargs="+", type=int, default=[8, 4, 2, 1])
    p.add_argument("--gamma", type=float, default=1e-6, help="Smoothing factor per voxel.")
    p.add_argument("--bspline-order", type=int, default=3)
    p.add_argument("--finite-difference-order", type=int, default=2)
    p.add_argument("--smooth", action="store_true", default=False,
                   help="Smooth images (not supported by this backend).")
    p.add_argument("--upsample-in-frequency-domain", action="store_true", default=False)
    p.add_argument("--threads", type=int, default=default_threads)
    p.add_argument("--save-likelihood-parameters", action="store_true", default=False)
    p.add_argument("--visualizer", action="store_true", default=False,
                   help="Turn visualizer on (not supported by this backend).")
    p.add_argument("--sampler", action="store_true", default=False, help="Enable sampler mode.")
    p.add_argument("--seed", type=int, default=12345)
    p.add_argument("--number-of-burnin", type=int, default=50)
    p.add_argument("--number-of-samples", type=int, default=50)
    p.add_argument("--sample-gamma", action="store_true", default=False)
    p.add_argument("--sample-gamma-every", type=int, default=1)
    p.add_argument("--do-not-sample-posteriors", action="store_true", default=False)
    p.add_argument("--do-not-sample-likelihood-parameters", action="store_true", default=False)
    p.add_argument("--do-not-sample-deformation-parameters", action="store_true", default=False)
    return p.parse_args(argv)


def _load_matrix(path):
    m = np.load(path) if path.endswith(".npy") else np.loadtxt(path)
    m = np.asarray(m, dtype=np.float64)
    if m.shape != (4, 4):
        raise ValueError(f"{path}: expected a 4x4 matrix, got {m.shape}")
    return m


def main(argv=None):
    from binder_amd import nifti, resample, utils
    from binder_amd.FastRegistration import FastRegistration

    args = parse_arguments(argv)
    os.makedirs(args.output, exist_ok=True)
    fixed_img, moving_img = nifti.load(args.fixed), nifti.load(args.moving)
    voxels, nodes = fixed_img.get_fdata(), moving_img.get_fdata()
    fixed_affine, moving_affine = fixed_img.affine, moving_img.affine
    mask_fixed = nifti.load(args.mask_fixed).get_fdata() if args.mask_fixed else None
    mask_moving = nifti.load(args.mask_moving).get_fdata() if args.mask_moving else None
    init = _load_matrix(args.initial_transformation) if args.initial_transformation else None

    registrator = FastRegistration(
        voxels=voxels, nodes=nodes, voxels_matrix=fixed_affine, nodes_matrix=moving_affine,
        mask_voxels=mask_fixed, mask_nodes=mask_moving, use_masks=args.automatic_mask,
        mask_threshold=args.mask_below, num_voxel_intesity_levels=args.bins_fixed,
        num_node_intensity_levels=args.bins_moving, bin_robust=not args.do_not_bin_robust,
        gamma=args.gamma, metric="MI", spline_order=args.bspline_order,
        topology_correction=not args.no_topology_correction,
        finite_difference_order=args.finite_difference_order, alpha=args.alpha,
        visualizer=args.visualizer, threads=args.threads, convergence_th=args.convergence_th,
        seed=args.seed, e1=args.e1, e2=args.e2, max_outer_it=args.max_outer_it,
        max_inner_it=args.max_inner_it, smooth=args.smooth,
        upsample_in_frequency_domain=args.upsample_in_frequency_domain,
        initial_transformation=init,
        initialize_with_center_of_mass=not args.do_not_initialize_center_of_mass,
        increase_gamma_at_DS_resolution=not args.do_not_increase_gamma)

    start = time.time()
    registrator.run(resolutions=args.resolutions, max_number_of_iterations=args.iterations)
    print("Time for EM: " + str(time.time() - start) + " seconds")

    if args.sampler:
        print("Running Gibbs sampler")
        start = time.time()
        mean_warp, cov_field = registrator.sample_param(
            args.number_of_burnin, args.number_of_samples, int(args.sample_gamma),
            args.sample_gamma_every, int(not args.do_not_sample_posteriors),
            int(not args.do_not_sample_likelihood_parameters),
            int(not args.do_not_sample_deformation_parameters))
        print("Time for Gibbs sampler: " + str(time.time() - start) + " seconds")
        if registrator.current_resolution != 1:
            r = registrator.current_resolution
            mean_warp = resample.resize(mean_warp - registrator.voxels_pos,
                                        registrator.orig_voxels_shape, order=1) * r
            mean_warp = mean_warp + registrator.full_voxels_pos
            cov_field = resample.resize(cov_field, registrator.orig_voxels_shape, order=1) * r

    out = args.output
    utils.save_warp(deformations=registrator.deformedVoxelPositions - registrator.voxels_pos,
                    outputDir=out, affine=fixed_affine)
    utils.save_image(generalImage=registrator.deformedVoxelPositions, outputDir=out,
                     affine=fixed_affine, outputName="deformation_field.nii.gz")
    if args.save_jacobian_determinant:
        from binder_amd._core import compute_jacobian_determinant_3D
        tmp = np.ascontiguousarray(np.asarray(registrator.deformedVoxelPositions, np.float64) /
                                   registrator.voxels_voxel_size[None, None, None, :])
        vs = registrator.voxels_voxel_size
        jac = compute_jacobian_determinant_3D(tmp.ravel(), tmp.shape[0], tmp.shape[1],
                                              tmp.shape[2], vs[0], vs[1], vs[2])
        utils.save_image(generalImage=jac, outputDir=out, affine=fixed_affine,
                         outputName="jacobian_determinant.nii.gz")
    if args.save_warped_image:
        utils.save_warped_image(deformed_position=registrator.deformedVoxelPositions,
                                movingImage=nodes, order=args.output_interpolation,
                                affine=fixed_affine, outputDir=out)
    if args.save_binned_images:
        utils.save_warped_image(deformed_position=registrator.deformedVoxelPositions,
                                movingImage=registrator.full_binned_nodes, order=0,
                                affine=fixed_affine, outputDir=out,
                                outputName="binned_warped.nii.gz")
        utils.save_image(generalImage=registrator.full_binned_voxels, outputDir=out,
                         affine=fixed_affine, outputName="binned_fixed.nii.gz")
    if args.sampler:
        utils.save_warp(deformations=mean_warp - registrator.voxels_pos, outputDir=out,
                        affine=fixed_affine, outputName="mean_warp_sample.nii.gz")
        utils.save_image(generalImage=mean_warp, outputDir=out, affine=fixed_affine,
                         outputName="mean_deformation_field_sample.nii.gz")
        utils.save_image(generalImage=cov_field, outputDir=out, affine=fixed_affine,
                         outputName="covariance_field_sample.nii.gz")
        if args.save_warped_image:
            utils.save_warped_image(deformed_position=mean_warp, movingImage=nodes,
                                    order=args.output_interpolation, affine=fixed_affine,
                                    outputDir=out, outputName="warped_sample.nii.gz")
    if args.save_masks:
        utils.save_image(generalImage=registrator.full_mask_voxels, outputDir=out,
                         affine=fixed_affine, outputName="mask_fixed.nii.gz")
        utils.save_image(generalImage=registrator.full_mask_nodes, outputDir=out,
                         affine=fixed_affine, outputName="mask_moving.nii.gz")
    if args.save_likelihood_parameters:
        utils.save_likelihood_parameters(registrator, out)
    print("Done! FASTMIRegistration exit WITHOUT ERRORS.")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
