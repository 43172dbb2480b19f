"""Multi-resolution driver of the registration path (the reference's
``BINDER.FastRegistration``, BINDER/FastRegistration.py:26-750), on the
gfx950 backend ``binder_amd._core``.

Same constructor keywords, ``run(resolutions, max_number_of_iterations)``,
``move_resolution(resolution)`` and ``sample_param(...)`` as the reference,
for the MI and SSD metrics with the non-linear (DCT-prior), affine or rigid
transformation.  The reference's feature likelihoods (MILikelihoodFilter,
SSDLikelihoodFilter) and spline-smoothed pyramids are not part of it:
asking for them raises NotImplementedError instead of silently running
something else.  nibabel and scikit-image are replaced by
``binder_amd.nifti`` and ``binder_amd.resample``.

Every registration and sampling call runs on the GPU through ``_core``;
only the pyramid bookkeeping (down-sampling the bins, resizing the field,
gamma and D per level) is host NumPy, as in the reference.
"""
import numpy as np

from . import _core as Registration
from . import nifti, utils


def _apply_affine(aff, pts):
    return nifti.apply_affine(aff, pts)


class FastRegistration:
    """Registers ``voxels`` (the fixed image) into ``nodes`` (the moving
    image).  See BINDER/FastRegistration.py:27-440 for the meaning of each
    keyword."""

    def __init__(self, voxels, nodes, voxels_matrix=np.eye(4), nodes_matrix=np.eye(4),
                 mask_voxels=None, mask_nodes=None, use_masks=False, mask_threshold=0,
                 num_voxel_intesity_levels=64, num_node_intensity_levels=64, alpha=2.0,
                 alpha_0=1, beta_0=1, gamma=5e-7, metric='MI', histogram_matching=False,
                 scale_shift=True, scale=1, shift=0, spline_order=3,
                 finite_difference_order=2, topology_correction=True, visualizer=False,
                 feature_moving=False, feature_fixed=False, precomputed_features_moving=None,
                 precomputed_features_fixed=None, number_of_filters_moving=1,
                 number_of_filters_fixed=1, fixed_features_reg=False, affine=False, rigid=False,
                 smooth=False, upsample_in_frequency_domain=False, convergence_th=5e-5,
                 e1=0.01, e2=100, max_outer_it=10, max_inner_it=10, threads=1, seed=12345,
                 debug=0, update_likelihood_parameters=1, initial_transformation=None,
                 initialize_with_center_of_mass=True, bin_robust=True,
                 include_image_as_feature=True, increase_gamma_at_DS_resolution=False):
        unsupported = []
        if metric not in ('MI', 'SSD'):
            unsupported.append(f"metric={metric!r}")
        if feature_moving or feature_fixed or precomputed_features_moving is not None \
                or precomputed_features_fixed is not None:
            unsupported.append("feature likelihoods")
        if smooth:
            unsupported.append("smooth=True (spline-smoothed pyramid)")
        if visualizer:
            unsupported.append("visualizer=True")
        if unsupported:
            raise NotImplementedError(
                "the gfx950 backend implements the MI and SSD metrics with the non-linear, "
                "affine and rigid transformations; not available: " + ", ".join(unsupported))
        if spline_order not in (1, 2, 3):
            raise ValueError("spline_order must be 1, 2 or 3")

        self.registratorObj = None
        self.metric = metric
        # affine wins over rigid, as in the reference's set-up
        # (FastRegistration.py:452-472)
        self.affine = bool(affine)
        self.rigid = bool(rigid) and not self.affine
        self.smooth = False
        self.visualizer = False
        self.feature_moving = False
        self.feature_fixed = False
        self.convergence_th = convergence_th
        self.debug = debug
        self.update_likelihood_parameters = update_likelihood_parameters
        self.upsample_in_frequency_domain = upsample_in_frequency_domain
        self.increase_gamma_at_DS_resolution = increase_gamma_at_DS_resolution
        self.threads = threads
        self.seed = seed
        self._D = None

        voxels = np.asarray(voxels, dtype=np.float64)
        nodes = np.asarray(nodes, dtype=np.float64)
        self.number_of_dimensions = nodes.ndim
        if self.number_of_dimensions == 2:
            voxels = voxels[:, :, None]
            nodes = nodes[:, :, None]

        if topology_correction:
            name = "topology_correction_2D" if self.number_of_dimensions == 2 else \
                "topology_correction_3D"
            if not hasattr(Registration, name):
                raise RuntimeError(
                    f"topology_correction=True was requested, but this backend has no {name}; "
                    "pass topology_correction=False")
            self.top_cor_obj = getattr(Registration, name)()
            self.e1, self.e2 = e1, e2
            self.max_inner_it, self.max_outer_it = max_inner_it, max_outer_it
        else:
            self.top_cor_obj = 0

        self.voxels = voxels
        self.nodes = nodes
        vm = np.asarray(voxels_matrix, dtype=np.float64)
        nm = np.asarray(nodes_matrix, dtype=np.float64)
        # voxel size = norm of each column of the linear part (FastRegistration.py:154-155)
        self.voxels_voxel_size = np.sqrt(np.sum(vm[0:3, 0:3] ** 2, axis=0))
        self.nodes_voxel_size = np.sqrt(np.sum(nm[0:3, 0:3] ** 2, axis=0))
        self.current_voxels_voxel_size = self.voxels_voxel_size
        self.current_nodes_voxel_size = self.nodes_voxel_size
        self.gamma_init = gamma

        if histogram_matching:
            nodes = utils.hist_norm(nodes, voxels)

        # masks: None stands for "all ones" (made on first read, see the
        # properties below); binning then skips the masking, which leaves
        # every value unchanged
        self._mask_shapes = (nodes.shape, voxels.shape)
        if mask_nodes is None:
            self._full_mask_nodes = utils.maskImage(nodes, th_value=mask_threshold) \
                if use_masks else None
        else:
            self._full_mask_nodes = np.asarray(mask_nodes).reshape(nodes.shape)
        if mask_voxels is None:
            self._full_mask_voxels = utils.maskImage(voxels, th_value=mask_threshold) \
                if use_masks else None
        else:
            self._full_mask_voxels = np.asarray(mask_voxels).reshape(voxels.shape)

        self.finite_difference_order = finite_difference_order
        self.orig_voxels_shape = voxels.shape
        self.orig_nodes_shape = nodes.shape
        self.num_voxel_intensity_levels = num_voxel_intesity_levels
        self.num_node_intensity_levels = num_node_intensity_levels

        def _bin(img, mask, levels):
            masked = img if mask is None else img * mask
            if bin_robust:
                b, _ = utils.bin_signal_robust(masked, levels)
            else:
                b = utils.bin_signal(masked, levels)
            if mask is not None:
                b[np.logical_not(mask)] = 0
            return b

        self.full_binned_voxels = _bin(voxels, self._full_mask_voxels,
                                       self.num_voxel_intensity_levels)
        self.full_binned_nodes = _bin(nodes, self._full_mask_nodes,
                                      self.num_node_intensity_levels)
        self.alpha = alpha
        # theta initialised to 1/L for all k, l (FastRegistration.py:224-225)
        self.theta = np.ones([self.num_node_intensity_levels, self.num_voxel_intensity_levels]) \
            / self.num_voxel_intensity_levels
        if metric == 'SSD':
            # sigma^2 starts at the mean squared bin difference over the
            # common extent of the two images (FastRegistration.py:241-259)
            tmp = np.minimum(self.full_binned_voxels.shape, self.full_binned_nodes.shape)
            sl = tuple(slice(0, int(n)) for n in tmp)
            diff = self.full_binned_nodes[sl].astype(np.float64) - \
                self.full_binned_voxels[sl].astype(np.float64)
            self.sigma_sq = float(np.sum(diff ** 2) / np.prod(tmp))
            self.alpha_0, self.beta_0 = alpha_0, beta_0
            if scale_shift:
                self.scale_shift, self.scale, self.shift = 1, scale, shift
            else:
                self.scale_shift, self.scale, self.shift = 0, 1, 0
        # starting linear transformation (FastRegistration.py:105-111): the
        # identity in the homogeneous (affine) or plain (rigid) form of the
        # image's dimensionality
        nd = self.number_of_dimensions
        if self.affine:
            self.A, self.t = np.eye(nd + 1), np.zeros(nd)
        if self.rigid:
            self.R, self.t = np.eye(nd), np.zeros(nd)

        # full-resolution positions: the grid indices, made on first read
        # (the GPU pyramid generates them on the device and never reads them)
        self._full_vpos = None
        self._full_npos = None
        self._deformed = None  # None: the field is still the initial positions
        self._dream = None

        # voxel positions expressed in node space (FastRegistration.py:375-395);
        # without a transformation they stay the grid indices at every level
        self._index_grid = initial_transformation is None and np.allclose(nm, vm)
        if initial_transformation is not None:
            self.full_voxels_pos = _apply_affine(initial_transformation, self.full_voxels_pos)
        elif not np.allclose(nm, vm):
            M = np.linalg.inv(nm) @ vm
            self.full_voxels_pos = _apply_affine(M, self.full_voxels_pos)
            if initialize_with_center_of_mass:
                from scipy.ndimage import center_of_mass
                com_nodes = np.array(center_of_mass(self.full_binned_nodes > 0))
                com_voxels = _apply_affine(M, np.array(center_of_mass(self.full_binned_voxels > 0)))
                self.full_voxels_pos = self.full_voxels_pos + (com_nodes - com_voxels)

        self.current_resolution = 1
        self._level_pos = {}
        self.binned_voxels = self.full_binned_voxels
        self.binned_nodes = self.full_binned_nodes
        self.nodes_shape = self.orig_nodes_shape
        self.voxels_shape = self.orig_voxels_shape
        self.non_zero_voxels = np.count_nonzero(self.binned_voxels)
        self._non_zero = {1: self.non_zero_voxels}  # per resolution (bins are fixed)
        self.spline_order = spline_order
        self._set_gammas(self.voxels_shape)
        # per-level record: (resolution, EM iterations, seconds)
        self.level_log = []

    def _grid(self, shape):
        g = np.stack(np.indices(shape, dtype=np.float64), axis=-1)
        return g[..., 0:2] if self.number_of_dimensions == 2 else g

    @property
    def full_voxels_pos(self):
        if self._full_vpos is None:
            self._full_vpos = self._grid(self.orig_voxels_shape)
        return self._full_vpos

    @full_voxels_pos.setter
    def full_voxels_pos(self, value):
        self._full_vpos = value

    @property
    def full_nodes_pos(self):
        if self._full_npos is None:
            self._full_npos = self._grid(self.orig_nodes_shape)
        return self._full_npos

    @full_nodes_pos.setter
    def full_nodes_pos(self, value):
        self._full_npos = value

    @property
    def deformedVoxelPositions(self):
        return self.full_voxels_pos if self._deformed is None else self._deformed

    @deformedVoxelPositions.setter
    def deformedVoxelPositions(self, value):
        self._deformed = value

    @property
    def dreamLocations(self):
        if self._dream is None:
            self._dream = np.zeros_like(self.deformedVoxelPositions)
        return self._dream

    @dreamLocations.setter
    def dreamLocations(self, value):
        self._dream = value

    @property
    def full_mask_voxels(self):
        if self._full_mask_voxels is None:
            self._full_mask_voxels = np.ones(self._mask_shapes[1])
        return self._full_mask_voxels

    @property
    def full_mask_nodes(self):
        if self._full_mask_nodes is None:
            self._full_mask_nodes = np.ones(self._mask_shapes[0])
        return self._full_mask_nodes

    # positions of the current level (full[::res] / res), made on first use:
    # the GPU pyramid generates the index grid on the device and never needs
    # them, and the strided copy costs tens of milliseconds at 160x192x160
    def _level_positions(self, which):
        res = self.current_resolution
        lazy = self._full_vpos is None if which == "v" else self._full_npos is None
        if res == 1:
            return self.full_voxels_pos if which == "v" else self.full_nodes_pos
        key = (which, res)
        if key not in self._level_pos:
            self._level_pos = {k: v for k, v in self._level_pos.items() if k[1] == res}
            if lazy and (which == "n" or self._index_grid):
                # full[::res] / res of the index grid is the level's index
                # grid exactly (k res / res == k in FP64): made at level size
                shape = self.voxels_shape if which == "v" else tuple(self.nodes_shape)
                self._level_pos[key] = self._grid(shape)
            else:
                full = self.full_voxels_pos if which == "v" else self.full_nodes_pos
                self._level_pos[key] = full[::res, ::res, ::res, :] / res
        return self._level_pos[key]

    @property
    def voxels_pos(self):
        return self._level_positions("v")

    @property
    def nodes_pos(self):
        return self._level_positions("n")

    @property
    def D(self):
        """The current level's DCT prior spectrum (utils.py:331-349); the GPU
        path makes it on the device instead and never reads this."""
        if self._D is None:
            if hasattr(Registration, "dct_prior_eigenvalues"):
                # the float32 values this backend keeps (utils.py:331-349 rounded)
                self._D = Registration.dct_prior_eigenvalues(
                    self.voxels_shape, self.finite_difference_order,
                    self.number_of_dimensions, voxel_size=self.current_voxels_voxel_size)
            else:
                self._D = utils.initialize_fourier_domain_smoothing(
                    self.voxels_shape, True, self.finite_difference_order,
                    self.number_of_dimensions, voxel_size=self.current_voxels_voxel_size)
        return self._D

    @D.setter
    def D(self, value):
        self._D = value

    def _set_gammas(self, shape_for_scaling):
        # gamma per voxel at 1 mm, scaled by voxel volume, voxel count and the
        # squared node size of each axis (FastRegistration.py:411-417)
        base = self.gamma_init * np.prod(self.current_voxels_voxel_size) * \
            np.prod(shape_for_scaling)
        self.gamma_x = base * self.current_nodes_voxel_size[0] ** 2
        self.gamma_y = base * self.current_nodes_voxel_size[1] ** 2
        self.gamma_z = base * self.current_nodes_voxel_size[2] ** 2

    def run(self, resolutions, max_number_of_iterations, level_iterations=None):
        """EM at each resolution of the pyramid, then (optionally) topology
        correction (FastRegistration.py:443-546).  `level_iterations` (not in
        the reference) replaces the per-level cap
        max_number_of_iterations * resolution with explicit counts."""
        for reg, cap in self.pyramid_levels(resolutions, max_number_of_iterations,
                                            level_iterations):
            reg.EM(cap, self.convergence_th,
                   update_likelihood_parameters=self.update_likelihood_parameters,
                   debug=self.debug)

    def pyramid_levels(self, resolutions, max_number_of_iterations, level_iterations=None):
        """The pyramid of run() as a generator: for each level it sets up the
        level's registration and yields (registration, iteration cap); the
        caller runs that level's EM (Registration.EM, or Batch.EM over several
        pairs' levels at once), and resuming the generator finishes the level
        (topology correction, theta hand-over) and sets up the next.  Not in
        the reference."""
        import time
        cls = Registration.Registration
        # with this backend the field stays on the GPU from level to level:
        # device resize into the next level's registration, device topology
        # correction; other backends (the oracle) take the host path
        chain = hasattr(cls, "resampled_from") and \
            (not self.top_cor_obj or hasattr(cls, "correct_topology"))
        prev = None  # the previous level's registration, whose field is on the GPU
        warm, self._warm = getattr(self, "_warm", None), None
        for li, resolution in enumerate(resolutions):
            res, cur = int(resolution), self.current_resolution
            # ... unless the frequency-domain up-sampling is asked for
            # (FastRegistration.py:566-580)
            on_gpu = chain and prev is not None and \
                not (self.upsample_in_frequency_domain and cur - res >= 0 and res != cur)
            start_warm = warm is not None and li == 0
            self.move_resolution(resolution, resize_field=not on_gpu and not start_warm)
            t0 = time.perf_counter()
            if start_warm:
                # the first level starts from the predicted field (run_warm)
                self.registratorObj = self._warm_registration(warm, res)
            elif on_gpu:
                # full[::res] / res of the index grid is the index grid again,
                # which the device generates
                vp = None if self._index_grid else np.ascontiguousarray(self.voxels_pos).ravel()
                sh = self.voxels_shape
                self.registratorObj = Registration.Registration.resampled_from(
                    prev, cur / res, vp, sh[0], sh[1], sh[2],
                    self.nodes_shape[0], self.nodes_shape[1], self.nodes_shape[2],
                    self.spline_order, self.non_zero_voxels)
                self._set_transformation(self.registratorObj)
            else:
                self.deformedVoxelPositions = np.ascontiguousarray(self.deformedVoxelPositions)
                self.registratorObj = self._new_registration(self.deformedVoxelPositions)
            self._set_likelihood(self.registratorObj)
            cap = max_number_of_iterations * self.current_resolution
            if level_iterations is not None:
                cap = int(level_iterations[li])
            yield self.registratorObj, cap
            if self.top_cor_obj and chain:
                # the same correction as correct_topology(), on the field where it lives
                n = 2 if self.number_of_dimensions == 2 else 3
                self.registratorObj.correct_topology(
                    self.e1, self.e2, self.max_outer_it, self.max_inner_it,
                    self.current_voxels_voxel_size[:n])
            last = li == len(resolutions) - 1
            nxt = None if last else int(resolutions[li + 1])
            if not chain or last or (self.upsample_in_frequency_domain and res - nxt >= 0
                                     and nxt != res):
                # the host copy is needed: end of the pyramid, a host-side
                # backend, or a frequency-domain up-sampling next
                self.deformedVoxelPositions = \
                    self.registratorObj.get_final_locations_as_numpy_array()
                if self.top_cor_obj and not chain:
                    self.deformedVoxelPositions = self.correct_topology(
                        self.deformedVoxelPositions)
            self._keep_likelihood_parameters(self.registratorObj)
            iters = len(self.registratorObj.get_objective_trace())
            prev = self.registratorObj
            self.level_log.append((resolution, iters, time.perf_counter() - t0))

    def _warm_registration(self, coef, res):
        """Registration at the current level whose field is the warm start
        predicted for this pair: the voxel positions plus the predicted
        low-frequency displacement (Registration.from_dct, evaluated on the
        GPU; a host evaluation for backends without it)."""
        cls = Registration.Registration
        sh = self.voxels_shape
        if hasattr(cls, "from_dct"):
            vp = None if self._index_grid else np.ascontiguousarray(self.voxels_pos).ravel()
            r = cls.from_dct(coef, self.orig_voxels_shape, res, vp, sh[0], sh[1], sh[2],
                             self.nodes_shape[0], self.nodes_shape[1], self.nodes_shape[2],
                             self.spline_order, self.non_zero_voxels)
            self._set_transformation(r)
            return r
        from .warmstart import coefficients_field
        start = coefficients_field(coef, self.orig_voxels_shape, res)
        start = start - np.stack(np.indices(sh, dtype=np.float64), -1) + self.voxels_pos
        return self._new_registration(start)

    def _start_objective(self, coef, res, iterations):
        """Objective (mlp) after `iterations` EM iterations at level `res`
        from the warm start (coef) or, for coef None, from the identity, and
        the likelihood parameters theta fitted there."""
        r = self._warm_registration(coef, res) if coef is not None else \
            self._new_registration(np.ascontiguousarray(self.voxels_pos))
        r.setMILikelihood(self.alpha, np.ascontiguousarray(self.theta).ravel(),
                          self.num_node_intensity_levels, self.num_voxel_intensity_levels,
                          self.threads, np.ascontiguousarray(self.binned_nodes).ravel(),
                          np.ascontiguousarray(self.binned_voxels).ravel())
        r.EM(int(iterations), -1e30, update_likelihood_parameters=self.update_likelihood_parameters,
             debug=0)
        theta = np.array(r.get_likelihood_parameters()).reshape(
            self.num_node_intensity_levels, self.num_voxel_intensity_levels)
        return float(r.get_objective_trace()[-1, 3]), theta

    def run_warm(self, net, resolutions=(1,), max_number_of_iterations=1000,
                 level_iterations=None, fallback_resolutions=(8, 4, 2, 1),
                 compare_iterations=2, compare_resolution=8):
        """Registration from a learned warm start (binder_amd.warmstart; not
        in the reference).  The network predicts the low-frequency field from
        this pair's half-resolution bins; the first level of `resolutions`
        starts from it instead of the identity, and the plain EM then runs
        to convergence at every level exactly as run() does.

        Fallback: the objective (mlp) after `compare_iterations` EM
        iterations at level `compare_resolution` is computed from the warm
        start and from the identity; if the warm start's is worse (higher),
        the plain cold pyramid `fallback_resolutions` runs instead
        (level_iterations then applies to it).  Otherwise the likelihood
        parameters theta fitted from the warm start in that check start the
        first level, as a coarser level's theta starts the next one in the
        plain pyramid.  Returns True when the warm start was used."""
        if self.number_of_dimensions != 3:
            raise ValueError("the learned warm start is defined for 3D registrations")
        if self.metric != 'MI' or self.affine or self.rigid:
            raise ValueError("the learned warm start is defined for the MI metric with the "
                             "non-linear transformation")
        coef = net.predict_pair(self.full_binned_voxels, self.full_binned_nodes,
                                self.num_voxel_intensity_levels, self.num_node_intensity_levels)
        self.warm_coef = coef
        res0 = int(resolutions[0])
        used = True
        if compare_iterations:
            cr = int(compare_resolution)
            self.move_resolution(cr, resize_field=False)
            (mw, theta_w), (mc, _) = (self._start_objective(coef, cr, compare_iterations),
                                      self._start_objective(None, cr, compare_iterations))
            self.warm_objectives = (mw, mc)
            used = mw <= mc
            if used:
                # like a coarser pyramid level, the check hands its fitted
                # likelihood parameters on to the first level run
                self.theta = np.ascontiguousarray(theta_w)
        self.warm_theta = np.ascontiguousarray(self.theta).copy()
        self.warm_start_used = used
        if not used:
            self.move_resolution(1, resize_field=False)
            self.run(fallback_resolutions, max_number_of_iterations,
                     level_iterations=level_iterations)
            return False
        self._warm = coef
        self.run(resolutions, max_number_of_iterations, level_iterations=level_iterations)
        return True

    def correct_topology(self, positions):
        """Topology correction of a field at the current resolution, as run()
        applies it after each level (FastRegistration.py:521-544)."""
        vs = self.current_voxels_voxel_size
        if self.number_of_dimensions == 2:
            f = np.ascontiguousarray(np.asarray(positions)[:, :, 0])
            out = self.top_cor_obj.correct_topology(
                f.ravel(), self.e1, self.e2, self.max_outer_it, self.max_inner_it,
                self.threads, vs, self.voxels_shape[0], self.voxels_shape[1])
            d = np.array(positions, dtype=np.float64)
            d[:, :, 0] = np.asarray(out).reshape(d[:, :, 0].shape)
            return d
        return self.top_cor_obj.correct_topology(
            np.ascontiguousarray(positions).ravel(), self.e1, self.e2, self.max_outer_it,
            self.max_inner_it, self.threads, vs, self.voxels_shape[0], self.voxels_shape[1],
            self.voxels_shape[2])

    def _new_registration(self, positions):
        vp = np.ascontiguousarray(self.voxels_pos)
        r = Registration.Registration(np.ascontiguousarray(positions).ravel(), self.threads,
                                      vp.ravel(), np.ascontiguousarray(self.nodes_pos).ravel(),
                                      vp.shape[0], vp.shape[1], vp.shape[2],
                                      self.nodes_shape[0], self.nodes_shape[1],
                                      self.nodes_shape[2], self.spline_order,
                                      self.non_zero_voxels)
        self._set_transformation(r)
        return r

    def _set_transformation(self, r):
        """The transformation model of the reference's set-up
        (FastRegistration.py:452-472): affine, rigid or the DCT prior."""
        if self.affine:
            r.setAffineTransformation(np.ascontiguousarray(self.A).ravel(),
                                      np.ascontiguousarray(self.t).ravel())
        elif self.rigid:
            r.setRigidTransformation(np.ascontiguousarray(self.R).ravel(),
                                     np.ascontiguousarray(self.t).ravel())
        else:
            self._set_nonlinear(r)

    def _set_likelihood(self, r):
        """The likelihood model of the reference's set-up
        (FastRegistration.py:474-488): MI with theta or SSD."""
        bn = np.ascontiguousarray(self.binned_nodes).ravel()
        bv = np.ascontiguousarray(self.binned_voxels).ravel()
        if self.metric == 'SSD':
            r.setSSDLikelihood(self.num_voxel_intensity_levels, self.threads, self.sigma_sq,
                               self.scale_shift, self.scale, self.shift, bn, bv,
                               self.alpha_0, self.beta_0)
        else:
            r.setMILikelihood(self.alpha, np.ascontiguousarray(self.theta).ravel(),
                              self.num_node_intensity_levels, self.num_voxel_intensity_levels,
                              self.threads, bn, bv)

    def _keep_likelihood_parameters(self, r):
        """What a level hands to the next (FastRegistration.py:516-524)."""
        if self.metric == 'SSD':
            # the reference unpacks get_likelihood_parameters() = [scale,
            # shift] as (shift, scale): the next level starts from the fitted
            # scale as its shift and the fitted shift as its scale.  Kept
            # as is, so that a multi-level SSD run matches the reference's.
            self.shift, self.scale = r.get_likelihood_parameters()
        else:
            self.theta = np.ascontiguousarray(np.array(
                r.get_likelihood_parameters()).reshape(
                    self.num_node_intensity_levels, self.num_voxel_intensity_levels))

    def _set_nonlinear(self, r):
        if hasattr(r, "setNonLinearTransformationDCTPrior") and \
                self.finite_difference_order in (1, 2):
            # the same D values (dct_prior_eigenvalues), made on the GPU
            r.setNonLinearTransformationDCTPrior(
                self.gamma_x, self.gamma_y, self.gamma_z, *self.current_voxels_voxel_size[:3],
                *self.current_nodes_voxel_size[:3], self.finite_difference_order)
            return
        self.D = np.ascontiguousarray(self.D)
        r.setNonLinearTransformation(self.gamma_x, self.gamma_y, self.gamma_z,
                                     self.current_voxels_voxel_size[0],
                                     self.current_voxels_voxel_size[1],
                                     self.current_voxels_voxel_size[2],
                                     self.current_nodes_voxel_size[0],
                                     self.current_nodes_voxel_size[1],
                                     self.current_nodes_voxel_size[2], self.D.ravel())

    def move_resolution(self, resolution, resize_field=True):
        """Down-sample the bins and positions by `resolution`, resample the
        current field to the new grid and rescale it, recompute D and gamma
        (FastRegistration.py:550-650).  resize_field=False (not in the
        reference) leaves the field to run(), which resizes it on the GPU."""
        res = int(resolution)
        self.current_voxels_voxel_size = self.voxels_voxel_size * res
        self.current_nodes_voxel_size = self.nodes_voxel_size * res
        self.voxels_shape = tuple(int(np.ceil(x / res)) for x in self.orig_voxels_shape)
        self.nodes_shape = np.array([int(np.ceil(x / res)) for x in self.orig_nodes_shape])
        cur = self.current_resolution
        # voxels_pos / nodes_pos now refer to level `res` (made on first use)
        self.current_resolution = res
        self._level_pos = {}
        if not resize_field:
            pass
        elif not self.upsample_in_frequency_domain or cur - res < 0:
            # skimage resize(order=1) semantics, on the GPU
            if self._index_grid and cur == 1 and self._deformed is None and \
                    self._full_vpos is None and hasattr(Registration, "resize_index_grid"):
                # the untouched initial field is the index grid: the GPU makes
                # it instead of receiving 24 bytes per voxel from the host
                self.deformedVoxelPositions = Registration.resize_index_grid(
                    self.orig_voxels_shape, self.number_of_dimensions, self.voxels_shape,
                    cur / res)
            else:
                self.deformedVoxelPositions = Registration.resize_field(
                    self.deformedVoxelPositions, self.voxels_shape, cur / res)
        elif res != cur:
            full = self.full_voxels_pos[::cur, ::cur, ::cur, :] / cur
            delta = np.ascontiguousarray(self.deformedVoxelPositions - full)
            sh = self.deformedVoxelPositions.shape
            tmp = Registration.NonLinearTransformation.up_sample(
                delta.ravel(), sh[0], sh[1], sh[2], self.voxels_shape[0], self.voxels_shape[1],
                self.voxels_shape[2], self.threads)
            self.deformedVoxelPositions = tmp * cur / res + self.voxels_pos

        self.binned_voxels = self.full_binned_voxels[::res, ::res, ::res]
        self.binned_nodes = self.full_binned_nodes[::res, ::res, ::res]
        self._D = None  # the level's D, made on first use (D below)
        if res not in self._non_zero:
            self._non_zero[res] = np.count_nonzero(self.binned_voxels)
        self.non_zero_voxels = self._non_zero[res]
        self._set_gammas(self.orig_voxels_shape if self.increase_gamma_at_DS_resolution
                         else self.voxels_shape)

    def sample_param(self, burnin, samples, sample_gamma=0, sample_gamma_every=1,
                     sample_posteriors=1, sample_likelihood_parameters=1,
                     sample_transformation_parameters=1):
        """Gibbs sampler from the current registration state; returns the
        mean field and the voxel-wise covariance (FastRegistration.py:652-746)."""
        self.deformedVoxelPositions = np.ascontiguousarray(self.deformedVoxelPositions)
        if sample_posteriors == 0:
            self.dreamLocations = np.ascontiguousarray(
                self.registratorObj.get_dream_locations())
        self.registratorObj = self._new_registration(self.deformedVoxelPositions)
        if sample_posteriors == 0:
            self.registratorObj.set_dream_locations(np.asarray(self.dreamLocations).ravel())
        self._set_likelihood(self.registratorObj)
        return self.registratorObj.sampler(self.seed, burnin, samples, sample_gamma,
                                           sample_gamma_every, sample_posteriors,
                                           sample_likelihood_parameters,
                                           sample_transformation_parameters)

    def visualize_output(self, positions):
        """The reference's matplotlib view is not part of this backend."""
        return None
