"""Python face of the gfx950 backend: the reference's ``BINDER._core`` API
(names, argument order and keywords of BINDER/src/gpu/bindings.cpp, plus the
CPU-only ``compute_jacobian_determinant_*`` of BINDER/src/cpu/bindings.cpp)
implemented as a ctypes shim over the C ABI in include/binder_c.h.

Arrays are accepted in any dtype/contiguity and converted like pybind11's
``c_style | forcecast``.  Every call runs on the GPU through
``libbinder_hip.so``; if the library is missing the import fails.
"""
import ctypes
import os

import numpy as np

backend = "hip"
has_topology_correction = True

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BINDER_HIP_LIB", os.path.join(_HERE, "libbinder_hip.so"))

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"BINDER HIP backend library not found at {LIB_PATH}; build it with "
        "`python binder_amd/build.py` (hipcc, gfx950)")

_lib = ctypes.CDLL(LIB_PATH)

_c_int = ctypes.c_int
_c_double = ctypes.c_double
_vp = ctypes.c_void_p
_P = ctypes.POINTER


def _sig(name, restype, *args):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(args)
    return f


_lib.binder_last_error.restype = ctypes.c_char_p
_lib.binder_backend.restype = ctypes.c_char_p

_reg_create = _sig("binder_registration_create", _c_int, _P(_vp), _vp, _c_int, _vp, _vp,
                   _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_double)
_reg_destroy = _sig("binder_registration_destroy", None, _vp)
_reg_set_mi = _sig("binder_registration_set_mi_likelihood", _c_int, _vp, _c_double, _vp,
                   _c_int, _c_int, _c_int, _vp, _vp)
_reg_set_ssd = _sig("binder_registration_set_ssd_likelihood", _c_int, _vp, _c_int, _c_int,
                    _c_double, _c_int, _c_double, _c_double, _vp, _vp, _c_double, _c_double)
_reg_get_ssd = _sig("binder_registration_get_ssd_parameters", _c_int, _vp, _vp)
_reg_set_affine = _sig("binder_registration_set_affine_transformation", _c_int, _vp, _vp, _vp)
_reg_set_rigid = _sig("binder_registration_set_rigid_transformation", _c_int, _vp, _vp, _vp)
_reg_get_linear = _sig("binder_registration_get_linear_parameters", _c_int, _vp, _vp,
                       _P(_c_int))
_reg_set_nlt = _sig("binder_registration_set_nonlinear_transformation", _c_int, _vp,
                    *([_c_double] * 9), _vp)
_reg_em = _sig("binder_registration_em", _c_int, _vp, _c_int, _c_double, _c_int, _c_int)
_reg_set_nlt_prior = _sig("binder_registration_set_nonlinear_transformation_dct_prior", _c_int,
                          _vp, *([_c_double] * 9), _c_int)
_reg_sampler = _sig("binder_registration_sampler", _c_int, _vp, *([_c_int] * 8), _vp, _vp)
_reg_sampler_run = _sig("binder_registration_sampler_run", _c_int, _vp, *([_c_int] * 8))
_reg_sampler_result = _sig("binder_registration_sampler_result", _c_int, _vp, _vp, _vp)
_reg_sampler_chains = _sig("binder_registration_sampler_chains", _c_int, _vp, _c_int, _vp,
                           *([_c_int] * 6))
_reg_get_final = _sig("binder_registration_get_final_locations", _c_int, _vp, _vp)
_reg_get_dream = _sig("binder_registration_get_dream_locations", _c_int, _vp, _vp)
_reg_set_dream = _sig("binder_registration_set_dream_locations", _c_int, _vp, _vp)
_reg_get_theta = _sig("binder_registration_get_likelihood_parameters", _c_int, _vp, _vp)
_reg_get_shape = _sig("binder_registration_get_shape", _c_int, _vp, _vp)
_reg_get_gamma = _sig("binder_registration_get_gamma", _c_int, _vp, _vp)
_reg_create_resampled = _sig("binder_registration_create_resampled", _c_int, _P(_vp), _vp,
                             _c_double, _vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                             _c_int, _c_double)
_reg_topo = _sig("binder_registration_correct_topology", _c_int, _vp, _c_double, _c_double,
                 _c_int, _c_int, _vp, _vp)
_reg_get_trace = _sig("binder_registration_get_objective_trace", _c_int, _vp, _vp, _P(_c_int))
_reg_e_step = _sig("binder_registration_e_step", _c_int, _vp, _P(_c_double), _vp)
_reg_set_prof = _sig("binder_registration_set_profiling", _c_int, _vp, _c_int)
_reg_get_times = _sig("binder_registration_get_kernel_times", _c_int, _vp, ctypes.c_char_p,
                      _c_int, _vp, _vp, _P(_c_int))
_nlt_create = _sig("binder_nlt_create", _c_int, _P(_vp), *([_c_double] * 9), _vp,
                   _c_int, _c_int, _c_int, _c_int, _c_double, _c_double)
_nlt_destroy = _sig("binder_nlt_destroy", None, _vp)
_nlt_smooth = _sig("binder_nlt_smooth_deformation", _c_int, _vp, _vp, _vp, _P(_c_double))
_nlt_sample = _sig("binder_nlt_sample_deformation_field", _c_int, _vp, _vp, _c_int,
                   ctypes.c_ulonglong, _vp)
_nlt_up = _sig("binder_nlt_up_sample", _c_int, _vp, *([_c_int] * 7), _vp)
_jac3 = _sig("binder_jacobian_determinant_3d", _c_int, _vp, _c_int, _c_int, _c_int,
             _c_double, _c_double, _c_double, _vp)
_jac2 = _sig("binder_jacobian_determinant_2d", _c_int, _vp, _c_int, _c_int, _c_double,
             _c_double, _vp)
_fold3 = _sig("binder_fold_mask_3d", _c_int, _vp, _c_int, _c_int, _c_int, _c_double,
              _c_double, _c_double, _c_double, _vp, _P(ctypes.c_longlong))
_topo3 = _sig("binder_topology_correction_3d", _c_int, _vp, _c_double, _c_double, _c_int,
              _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _vp, _vp)


_philox = _sig("binder_philox4x32_10", _c_int, _vp, _vp, _vp)
_reg_set_dct = _sig("binder_registration_set_field_from_dct", _c_int, _vp, _vp, _c_int, _c_int,
                    _c_int, _c_int, _c_int)
_reg_create_dct = _sig("binder_registration_create_from_dct", _c_int, _P(_vp), _vp, _c_int,
                       _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _c_int, _c_int, _c_int,
                       _c_int, _c_int, _c_int, _c_double)


def _dct_coefficients(coef):
    """coef as float32 [3, C, C, C] (C in 1..16) and C."""
    c = _arr(coef, np.float32)
    C = int(round((c.size / 3) ** (1.0 / 3.0)))
    if C < 1 or C > 16 or c.size != 3 * C ** 3:
        raise ValueError("coef must hold 3 x C x C x C values, C in 1..16")
    return c, C


_dct_prior = _sig("binder_dct_prior_eigenvalues", _c_int, _c_int, _c_int, _c_int, _c_int,
                  _c_int, _vp, _vp)


def dct_prior_eigenvalues(size, finite_difference_order=2, number_of_dimensions=3,
                          voxel_size=(1.0, 1.0, 1.0)):
    """float32 values of utils.initialize_fourier_domain_smoothing(size, True,
    order, nd, voxel_size=...) (BINDER/utils.py:331-349), computed by a host
    loop over the separable per-axis terms."""
    nd = int(number_of_dimensions)
    dims = [int(n) for n in list(size)[:nd]]
    vs = np.ascontiguousarray(np.resize(np.asarray(voxel_size, np.float64), 3))
    out = np.empty(dims, np.float32)
    full = dims + [1] * (3 - nd)
    _check(_dct_prior(full[0], full[1], full[2], nd, int(finite_difference_order), _ptr(vs),
                      _ptr(out)))
    return out


def philox4x32_10(ctr, key):
    """Philox4x32-10 block (host-side) used by the sampler's random streams."""
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    _check(_philox(_ptr(c), _ptr(k), _ptr(out)))
    return out


def _check(rc):
    if rc == 0:
        return
    msg = _lib.binder_last_error().decode(errors="replace")
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def _arr(a, dtype):
    """pybind11-style c_style|forcecast conversion."""
    return np.ascontiguousarray(np.asarray(a), dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(_vp)


class Registration:
    """Registration(final_locations, threads, voxel_pos, node_pos, Nx_v, Ny_v,
    Nz_v, Nx_n, Ny_n, Nz_n, spline_order, non_zero_voxels)

    GPU-resident EM registration and Gibbs sampler
    (BINDER/src/gpu/bindings.cpp:17-157, BINDER/src/cpu/Registration.cpp)."""

    def __init__(self, final_locations, threads, voxel_pos, node_pos, Nx_v, Ny_v, Nz_v,
                 Nx_n, Ny_n, Nz_n, spline_order, non_zero_voxels):
        self._h = None
        fl = _arr(final_locations, np.float32)
        vp = _arr(voxel_pos, np.float32)
        N = 2 if int(Nz_v) == 1 else 3
        n = int(Nx_v) * int(Ny_v) * int(Nz_v) * N
        if fl.size != n or vp.size != n:
            raise ValueError(f"final_locations/voxel_pos must have {n} values")
        h = _vp()
        _check(_reg_create(ctypes.byref(h), _ptr(fl), int(threads), _ptr(vp), None,
                           int(Nx_v), int(Ny_v), int(Nz_v), int(Nx_n), int(Ny_n), int(Nz_n),
                           int(spline_order), float(non_zero_voxels)))
        self._h = h
        self._shape = (int(Nx_v), int(Ny_v), int(Nz_v), N)
        self._nn = int(Nx_n) * int(Ny_n) * int(Nz_n)
        self._KL = None
        self._ssd = False

    @classmethod
    def resampled_from(cls, src, scale, voxel_pos, Nx_v, Ny_v, Nz_v, Nx_n, Ny_n, Nz_n,
                       spline_order, non_zero_voxels):
        """A Registration on a new grid whose field is `src`'s current field
        resized on the GPU (the pyramid's resize, times `scale`) -- what
        Registration(resize(src.get_final_locations_as_numpy_array()), ...)
        computes, without the host round trip.  voxel_pos=None stands for
        the grid indices."""
        self = cls.__new__(cls)
        self._h = None
        N = 2 if int(Nz_v) == 1 else 3
        vp = None
        if voxel_pos is not None:  # None: the grid indices, generated on the GPU
            vp = _arr(voxel_pos, np.float32)
            if vp.size != int(Nx_v) * int(Ny_v) * int(Nz_v) * N:
                raise ValueError("voxel_pos has the wrong size")
        h = _vp()
        _check(_reg_create_resampled(ctypes.byref(h), src._h, float(scale),
                                     None if vp is None else _ptr(vp),
                                     int(Nx_v), int(Ny_v), int(Nz_v), int(Nx_n), int(Ny_n),
                                     int(Nz_n), int(spline_order), float(non_zero_voxels)))
        self._h = h
        self._shape = (int(Nx_v), int(Ny_v), int(Nz_v), N)
        self._nn = int(Nx_n) * int(Ny_n) * int(Nz_n)
        self._KL = None
        self._ssd = False
        return self

    @classmethod
    def from_dct(cls, coef, full_shape, resolution, voxel_pos, Nx_v, Ny_v, Nz_v, Nx_n, Ny_n,
                 Nz_n, spline_order, non_zero_voxels):
        """A Registration whose field starts at the learned warm start of a
        pyramid level (see set_field_from_dct), made on the GPU without a
        field upload.  voxel_pos=None stands for the grid indices.  No
        reference counterpart."""
        self = cls.__new__(cls)
        self._h = None
        c, C = _dct_coefficients(coef)
        vp = None
        if voxel_pos is not None:
            vp = _arr(voxel_pos, np.float32)
            if vp.size != int(Nx_v) * int(Ny_v) * int(Nz_v) * 3:
                raise ValueError("voxel_pos has the wrong size")
        fx, fy, fz = (int(n) for n in full_shape)
        h = _vp()
        _check(_reg_create_dct(ctypes.byref(h), _ptr(c), C, fx, fy, fz, int(resolution),
                               None if vp is None else _ptr(vp), int(Nx_v), int(Ny_v),
                               int(Nz_v), int(Nx_n), int(Ny_n), int(Nz_n), int(spline_order),
                               float(non_zero_voxels)))
        self._h = h
        self._shape = (int(Nx_v), int(Ny_v), int(Nz_v), 3)
        self._nn = int(Nx_n) * int(Ny_n) * int(Nz_n)
        self._KL = None
        self._ssd = False
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            _reg_destroy(self._h)
            self._h = None

    def setMILikelihood(self, alpha, theta, K, L, threads, binned_nodes, binned_voxels):
        th = _arr(theta, np.float64)
        bn = _arr(binned_nodes, np.int32)
        bv = _arr(binned_voxels, np.int32)
        if th.size != int(K) * int(L):
            raise ValueError("theta must have K*L values")
        if bn.size != self._nn or bv.size != int(np.prod(self._shape[:3])):
            raise ValueError("binned_nodes / binned_voxels have the wrong size")
        _check(_reg_set_mi(self._h, float(alpha), _ptr(th), int(K), int(L), int(threads),
                           _ptr(bn), _ptr(bv)))
        self._KL = (int(K), int(L))
        self._ssd = False

    def setSSDLikelihood(self, bins, threads, sigma_sq, scale_shift, scale, shift,
                         binned_nodes, binned_voxels, alpha_0, beta_0):
        """Sum-of-squared-differences likelihood (BINDER/src/cpu/bindings.cpp:120-152,
        SSDLikelihood.cpp): both images binned to [0, bins), bins <= 64;
        with scale_shift the EM fits shift and scale, and sigma_sq always."""
        bn = _arr(binned_nodes, np.int32)
        bv = _arr(binned_voxels, np.int32)
        if bn.size != self._nn or bv.size != int(np.prod(self._shape[:3])):
            raise ValueError("binned_nodes / binned_voxels have the wrong size")
        _check(_reg_set_ssd(self._h, int(bins), int(threads), float(sigma_sq),
                            int(bool(scale_shift)), float(scale), float(shift), _ptr(bn),
                            _ptr(bv), float(alpha_0), float(beta_0)))
        self._KL = (int(bins), int(bins))
        self._ssd = True

    def get_ssd_parameters(self):
        """[scale, shift, sigma_sq] of the SSD likelihood.  No reference
        counterpart (the reference prints them)."""
        out = np.empty(3, np.float64)
        _check(_reg_get_ssd(self._h, _ptr(out)))
        return out.tolist()

    def _linear_matrix(self, M, homogeneous):
        N = self._shape[3]
        m = np.asarray(M, np.float64).ravel()
        if homogeneous and m.size == N * N:
            # the N x N linear part: the homogeneous row and column are the
            # identity's (FastRegistration passes np.eye(3) as the start)
            h = np.eye(N + 1)
            h[:N, :N] = m.reshape(N, N)
            m = h.ravel()
        want = (N + 1) ** 2 if homogeneous else N * N
        if m.size != want:
            raise ValueError(f"the matrix must have {want} values")
        return np.ascontiguousarray(m)

    def setAffineTransformation(self, A, t):
        """Affine transformation model (BINDER/src/cpu/bindings.cpp:223-235,
        AffineTransformation.cpp): A (N+1) x (N+1) homogeneous or its N x N
        linear part, t of N values; the EM's least-squares fit replaces both."""
        a = self._linear_matrix(A, True)
        tv = _arr(t, np.float64)
        if tv.size != self._shape[3]:
            raise ValueError("t must have N values")
        _check(_reg_set_affine(self._h, _ptr(a), _ptr(tv)))

    def setRigidTransformation(self, R, t):
        """Rigid transformation model (BINDER/src/cpu/bindings.cpp:236-247,
        RigidTransformation.cpp): R N x N, t of N values."""
        r = self._linear_matrix(R, False)
        tv = _arr(t, np.float64)
        if tv.size != self._shape[3]:
            raise ValueError("t must have N values")
        _check(_reg_set_rigid(self._h, _ptr(r), _ptr(tv)))

    def get_linear_parameters(self):
        """The affine ((N+1) x (N+1)) or rigid (N x N) matrix as last fitted,
        row-major, followed by t.  No reference counterpart (the reference
        prints them)."""
        out = np.empty(19, np.float64)
        n = _c_int(0)
        _check(_reg_get_linear(self._h, _ptr(out), ctypes.byref(n)))
        return out[:n.value].tolist()

    def setNonLinearTransformation(self, gamma_x, gamma_y, gamma_z, voxels_size_x,
                                   voxels_size_y, voxels_size_z, nodes_size_x, nodes_size_y,
                                   nodes_size_z, D):
        Dv = _arr(D, np.float32)
        if Dv.size != int(np.prod(self._shape[:3])):
            raise ValueError("D must have Nx_v*Ny_v*Nz_v values")
        _check(_reg_set_nlt(self._h, float(gamma_x), float(gamma_y), float(gamma_z),
                            float(voxels_size_x), float(voxels_size_y), float(voxels_size_z),
                            float(nodes_size_x), float(nodes_size_y), float(nodes_size_z),
                            _ptr(Dv)))

    def setNonLinearTransformationDCTPrior(self, gamma_x, gamma_y, gamma_z, voxels_size_x,
                                           voxels_size_y, voxels_size_z, nodes_size_x,
                                           nodes_size_y, nodes_size_z,
                                           finite_difference_order=2):
        """setNonLinearTransformation with D = dct_prior_eigenvalues(grid,
        order, voxel sizes) made on the GPU (the same float32 values; no D
        upload).  No reference counterpart."""
        _check(_reg_set_nlt_prior(self._h, float(gamma_x), float(gamma_y), float(gamma_z),
                                  float(voxels_size_x), float(voxels_size_y),
                                  float(voxels_size_z), float(nodes_size_x),
                                  float(nodes_size_y), float(nodes_size_z),
                                  int(finite_difference_order)))

    def EM(self, max_EM_iterations, convergence_th, update_likelihood_parameters, debug):
        _check(_reg_em(self._h, int(max_EM_iterations), float(convergence_th),
                       int(bool(update_likelihood_parameters)), int(bool(debug))))

    def sampler(self, seed, N_b, N_s, sample_gamma, sample_gamma_every, sample_posteriors,
                sample_likelihood_parameters, sample_transformation_parameters):
        Nx, Ny, Nz, N = self._shape
        mean = np.empty((Nx, Ny, Nz, N), np.float64)
        cov = np.empty((Nx, Ny, Nz, N, N), np.float64)
        _check(_reg_sampler(self._h, int(seed), int(N_b), int(N_s), int(sample_gamma),
                            int(sample_gamma_every), int(sample_posteriors),
                            int(sample_likelihood_parameters),
                            int(sample_transformation_parameters), _ptr(mean), _ptr(cov)))
        return mean, cov

    def sampler_run(self, seed, N_b, N_s, sample_gamma=0, sample_gamma_every=1,
                    sample_posteriors=1, sample_likelihood_parameters=1,
                    sample_transformation_parameters=1):
        """Run the sweeps, keeping mean/M2 on the device (see sampler_result)."""
        _check(_reg_sampler_run(self._h, int(seed), int(N_b), int(N_s), int(sample_gamma),
                                int(sample_gamma_every), int(sample_posteriors),
                                int(sample_likelihood_parameters),
                                int(sample_transformation_parameters)))

    def sampler_chains(self, seeds, N_b, N_s, sample_posteriors=1,
                       sample_likelihood_parameters=1, sample_transformation_parameters=1,
                       use_graph=True):
        """Independent chains (one per seed) from the current state, pooled
        into one mean/M2 kept on the device (see sampler_result); with
        use_graph the sweeps of chains 1.. replay one captured HIP graph.
        No reference counterpart (the reference's sampler() is one chain)."""
        sd = np.ascontiguousarray(seeds, np.int32)
        _check(_reg_sampler_chains(self._h, int(sd.size), _ptr(sd), int(N_b), int(N_s),
                                   int(sample_posteriors), int(sample_likelihood_parameters),
                                   int(sample_transformation_parameters), int(bool(use_graph))))

    def sampler_result(self):
        Nx, Ny, Nz, N = self._shape
        mean = np.empty((Nx, Ny, Nz, N), np.float64)
        cov = np.empty((Nx, Ny, Nz, N, N), np.float64)
        _check(_reg_sampler_result(self._h, _ptr(mean), _ptr(cov)))
        return mean, cov

    def set_field_from_dct(self, coef, full_shape, resolution):
        """Sets the current field to the warm start of a pyramid level at
        `resolution`: the displacement whose C x C x C lowest orthonormal
        DCT-II coefficients per component (on the full-resolution grid
        `full_shape`) are sqrt(prod(full_shape)) * coef[3,C,C,C], sampled at
        the full-resolution voxels resolution * p and divided by the
        resolution.  No reference counterpart (see binder_amd.warmstart)."""
        c, C = _dct_coefficients(coef)
        fx, fy, fz = (int(n) for n in full_shape)
        _check(_reg_set_dct(self._h, _ptr(c), C, fx, fy, fz, int(resolution)))

    def get_final_locations_as_numpy_array(self):
        out = np.empty(self._shape, np.float32)
        _check(_reg_get_final(self._h, _ptr(out)))
        return out

    def get_dream_locations(self):
        out = np.empty(self._shape, np.float32)
        _check(_reg_get_dream(self._h, _ptr(out)))
        return out.ravel().tolist()

    def set_dream_locations(self, dream_locations):
        d = _arr(dream_locations, np.float32)
        if d.size != int(np.prod(self._shape)):
            raise ValueError("dream_locations has the wrong size")
        _check(_reg_set_dream(self._h, _ptr(d)))

    def get_likelihood_parameters(self):
        if self._KL is None:
            raise RuntimeError("setMILikelihood was not called")
        # SSD: [scale, shift] (SSDLikelihood.cpp:328-334); MI: theta (K x L)
        out = np.empty(2 if self._ssd else self._KL[0] * self._KL[1], np.float64)
        _check(_reg_get_theta(self._h, _ptr(out)))
        return out.tolist()

    def get_shape(self):
        return list(self._shape)

    # -- additions beyond the reference surface ----------------------------
    def get_gamma(self):
        """Current deformation-prior weights [gamma_x, gamma_y, gamma_z]: as
        set, or as left by the last nu draw of a sampler call with
        sample_gamma (NonLinearTransformation.cpp:116-124, :140-148)."""
        out = np.empty(3, np.float64)
        _check(_reg_get_gamma(self._h, _ptr(out)))
        return out.tolist()

    def get_objective_trace(self):
        """Rows of (mll, prior_theta, prior_d, mlp) of the last EM call."""
        rows = _c_int(0)
        _check(_reg_get_trace(self._h, None, ctypes.byref(rows)))
        out = np.empty((max(rows.value, 1), 4), np.float64)
        _check(_reg_get_trace(self._h, _ptr(out), ctypes.byref(rows)))
        return out[:rows.value]

    def e_step(self, return_histogram=False):
        """One E-step; returns mll (and the soft joint histogram)."""
        mll = _c_double(0.0)
        hist = None
        if return_histogram:
            hist = np.empty(self._KL, np.float64)
        _check(_reg_e_step(self._h, ctypes.byref(mll), _ptr(hist) if hist is not None else None))
        return (mll.value, hist) if return_histogram else mll.value

    def correct_topology(self, e1, e2, max_outer_it, max_it, voxel_sizes):
        """Topology correction of the current field in place on the GPU (the
        computation of topology_correction_3D / _2D.correct_topology without
        the host round trip); returns the stats of the last call."""
        vs = _arr(voxel_sizes, np.float64)
        if vs.size < self._shape[3]:
            raise ValueError("voxel_sizes needs one value per component")
        stats = np.zeros(6, np.float64)
        _check(_reg_topo(self._h, float(e1), float(e2), int(max_outer_it), int(max_it),
                         _ptr(vs), _ptr(stats)))
        return stats

    def set_profiling(self, enable):
        _check(_reg_set_prof(self._h, int(bool(enable))))

    def get_kernel_times(self):
        """{kernel name: (total ms, launches)} recorded since set_profiling(True)."""
        cap = 64
        names = ctypes.create_string_buffer(4096)
        ms = np.zeros(cap, np.float64)
        cnt = np.zeros(cap, np.int64)
        n = _c_int(cap)
        _check(_reg_get_times(self._h, names, 4096, _ptr(ms), _ptr(cnt), ctypes.byref(n)))
        keys = names.value.decode().split("\n")[: n.value]
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(keys)}


_batch_create = _sig("binder_batch_create", _c_int, _P(_vp), _c_int, _vp)
_batch_destroy = _sig("binder_batch_destroy", None, _vp)
_batch_em = _sig("binder_batch_em", _c_int, _vp, _vp, _c_double, _c_int, _vp)
_batch_set = _sig("binder_batch_set", _c_int, _vp, _c_int, _vp, _c_int, _c_int)
_batch_run = _sig("binder_batch_run", _c_int, _vp, _c_double, _c_int, _vp)


class Batch:
    """Slots of Registration objects whose EM calls run in lockstep: each
    iteration runs every running slot's E-step in one launch, their
    histogram reductions and theta steps in one launch each, their
    smoothings as batched launches (small grids) or on a few streams (large
    grids), with one host read of all objectives.  Each registration ends
    exactly where its own EM(...) call would.  Registrations must share K
    and L (K <= 32) and dimensionality, and outlive their slot.  No
    reference counterpart (the reference registers one pair per object,
    BINDER/FastRegistration.py:443-450).

    Batch(registrations) fills one slot per registration; Batch(n) makes n
    empty slots.  EM() runs one EM call per slot to the end; set() and run()
    let slots finish and be refilled (with another pair, or the next
    pyramid level of the same pair) while the others keep iterating."""

    def __init__(self, registrations):
        self._h = None
        if isinstance(registrations, int):
            self._regs = [None] * registrations
        else:
            self._regs = list(registrations)
        arr = (_vp * len(self._regs))(*[r._h.value if r is not None else None
                                        for r in self._regs])
        h = _vp()
        _check(_batch_create(ctypes.byref(h), len(self._regs), arr))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _batch_destroy(self._h)
            self._h = None

    def __len__(self):
        return len(self._regs)

    def EM(self, max_EM_iterations, convergence_th, update_likelihood_parameters=1):
        """EM of every slot's registration (Registration.EM semantics per
        registration; max_EM_iterations is one cap or one per slot).
        Returns the iterations each ran."""
        n = len(self._regs)
        caps = np.ascontiguousarray(np.broadcast_to(np.asarray(max_EM_iterations), (n,)),
                                    np.int32)
        out = np.zeros(n, np.int32)
        _check(_batch_em(self._h, _ptr(caps), float(convergence_th),
                         int(bool(update_likelihood_parameters)), _ptr(out)))
        return out

    def set(self, slot, registration, max_EM_iterations, update_likelihood_parameters=1):
        """Puts `registration` (None: nothing) into `slot` with an EM call of
        max_EM_iterations ahead of it; it starts at the next run()."""
        _check(_batch_set(self._h, int(slot), None if registration is None else
                          registration._h, int(max_EM_iterations),
                          int(bool(update_likelihood_parameters))))
        self._regs[int(slot)] = registration

    def run(self, convergence_th, min_finished=1):
        """Iterates every running slot until at least `min_finished` of them
        finish their EM call (or all have).  Returns, per slot, the
        iterations of a call that finished during this run, else -1."""
        out = np.full(len(self._regs), -1, np.int32)
        _check(_batch_run(self._h, float(convergence_th), int(min_finished), _ptr(out)))
        return out


class NonLinearTransformation:
    """NonLinearTransformation(gamma_x, gamma_y, gamma_z, voxels_size_x..z,
    nodes_size_x..z, D, Nx_v, Ny_v, Nz_v, threads, sigma_spline, spline_offset)
    (BINDER/src/gpu/bindings.cpp:159-255)."""

    def __init__(self, gamma_x, gamma_y, gamma_z, voxels_size_x, voxels_size_y,
                 voxels_size_z, nodes_size_x, nodes_size_y, nodes_size_z, D, Nx_v, Ny_v,
                 Nz_v, threads, sigma_spline, spline_offset):
        self._h = None
        Dv = _arr(D, np.float32)
        if Dv.size != int(Nx_v) * int(Ny_v) * int(Nz_v):
            raise ValueError("D must have Nx_v*Ny_v*Nz_v values")
        h = _vp()
        _check(_nlt_create(ctypes.byref(h), float(gamma_x), float(gamma_y), float(gamma_z),
                           float(voxels_size_x), float(voxels_size_y), float(voxels_size_z),
                           float(nodes_size_x), float(nodes_size_y), float(nodes_size_z),
                           _ptr(Dv), int(Nx_v), int(Ny_v), int(Nz_v), int(threads),
                           float(sigma_spline), float(spline_offset)))
        self._h = h
        self._shape = (int(Nx_v), int(Ny_v), int(Nz_v), 2 if int(Nz_v) == 1 else 3)
        self.seed = 1234

    def __del__(self):
        if getattr(self, "_h", None):
            _nlt_destroy(self._h)
            self._h = None

    def get_shape(self):
        return list(self._shape)

    def smooth_deformation(self, deformation):
        d = _arr(deformation, np.float32)
        if d.size != int(np.prod(self._shape)):
            raise ValueError("deformation has the wrong size")
        out = np.empty(self._shape, np.float32)
        cost = _c_double(0.0)
        _check(_nlt_smooth(self._h, _ptr(d), _ptr(out), ctypes.byref(cost)))
        return cost.value, out

    def sampleDeformationField(self, deformationField, number_of_samples):
        d = _arr(deformationField, np.float32)
        if d.size != int(np.prod(self._shape)):
            raise ValueError("deformationField has the wrong size")
        S = int(number_of_samples)
        out = np.empty((S,) + self._shape, np.float32)
        _check(_nlt_sample(self._h, _ptr(d), S, int(self.seed), _ptr(out)))
        self.seed += 1
        return out

    @staticmethod
    def up_sample(deformation_field, N_x, N_y, N_z, N_x_o, N_y_o, N_z_o, threads):
        f = _arr(deformation_field, np.float64)
        N = 2 if int(N_z_o) == 1 else 3
        out = np.empty((int(N_x_o), int(N_y_o), int(N_z_o), N), np.float64)
        _check(_nlt_up(_ptr(f), int(N_x), int(N_y), int(N_z), int(N_x_o), int(N_y_o),
                       int(N_z_o), int(threads), _ptr(out)))
        return out


class topology_correction_3D:
    """topology_correction_3D().correct_topology(h_init, e1, e2, max_outer_it,
    max_it, threads, voxel_sizes, Nx_v, Ny_v, Nz_v)
    (BINDER/src/gpu/bindings.cpp:257-287)."""

    def __init__(self):
        self.last_stats = None

    def correct_topology(self, h_init, e1, e2, max_outer_it, max_it, threads, voxel_sizes,
                         Nx_v, Ny_v, Nz_v):
        h = _arr(h_init, np.float64)
        vs = _arr(voxel_sizes, np.float64)
        if h.size != int(Nx_v) * int(Ny_v) * int(Nz_v) * 3 or vs.size < 3:
            raise ValueError("h_init / voxel_sizes have the wrong size")
        out = np.empty((int(Nx_v), int(Ny_v), int(Nz_v), 3), np.float64)
        stats = np.zeros(6, np.float64)
        _check(_topo3(_ptr(h), float(e1), float(e2), int(max_outer_it), int(max_it),
                      int(threads), _ptr(vs), int(Nx_v), int(Ny_v), int(Nz_v), _ptr(out),
                      _ptr(stats)))
        self.last_stats = stats
        return out


_splw = _sig("binder_spline_weights", _c_int, _c_int, _c_double, _c_int, _vp)


class Spline:
    """Spline(order, is2D) with .spline(d) and .log_spline(d)
    (BINDER/src/cpu/bindings.cpp:436-445): the 4 B-spline weights (or their
    logs) of order 1, 2 or 3 at fractional distance d, as a list."""

    def __init__(self, order, is2D):
        if int(order) not in (1, 2, 3):
            raise ValueError("spline order must be 1, 2 or 3")
        self.order = int(order)
        self.is2D = bool(is2D)

    def _eval(self, d, log_space):
        out = np.empty(4, np.float64)
        _check(_splw(self.order, float(d), int(log_space), _ptr(out)))
        return out.tolist()

    def spline(self, d):
        return self._eval(d, 0)

    def log_spline(self, d):
        return self._eval(d, 1)


_topo2 = _sig("binder_topology_correction_2d", _c_int, _vp, _c_double, _c_double, _c_int,
              _c_int, _c_int, _vp, _c_int, _c_int, _vp, _vp)


class topology_correction_2D:
    """topology_correction_2D().correct_topology(h_init, e1, e2, max_outer_it,
    max_it, threads, voxel_sizes, Nx_v, Ny_v)
    (BINDER/src/cpu/bindings.cpp:449-478); FP64 on the GPU."""

    def __init__(self):
        self.last_stats = None

    def correct_topology(self, h_init, e1, e2, max_outer_it, max_it, threads, voxel_sizes,
                         Nx_v, Ny_v):
        h = _arr(h_init, np.float64)
        vs = _arr(voxel_sizes, np.float64)
        if h.size != int(Nx_v) * int(Ny_v) * 2 or vs.size < 2:
            raise ValueError("h_init / voxel_sizes have the wrong size")
        out = np.empty((int(Nx_v), int(Ny_v), 2), np.float64)
        stats = np.zeros(6, np.float64)
        _check(_topo2(_ptr(h), float(e1), float(e2), int(max_outer_it), int(max_it),
                      int(threads), _ptr(vs), int(Nx_v), int(Ny_v), _ptr(out), _ptr(stats)))
        self.last_stats = stats
        return out


def compute_jacobian_determinant_3D(field, Nx, Ny, Nz, voxel_size_x, voxel_size_y,
                                    voxel_size_z):
    """BINDER/src/cpu/bindings.cpp:564-593; FP64, bit-exact with the CPU."""
    f = _arr(field, np.float64)
    if f.size != int(Nx) * int(Ny) * int(Nz) * 3:
        raise ValueError("field must have Nx*Ny*Nz*3 values")
    out = np.empty((int(Nx), int(Ny), int(Nz)), np.float64)
    _check(_jac3(_ptr(f), int(Nx), int(Ny), int(Nz), float(voxel_size_x),
                 float(voxel_size_y), float(voxel_size_z), _ptr(out)))
    return out


def compute_jacobian_determinant_2D(field, Nx, Ny, voxel_size_x, voxel_size_y):
    """BINDER/src/cpu/bindings.cpp:530-562; FP64, bit-exact with the CPU."""
    f = _arr(field, np.float64)
    if f.size != int(Nx) * int(Ny) * 2:
        raise ValueError("field must have Nx*Ny*2 values")
    out = np.empty((int(Nx), int(Ny)), np.float64)
    _check(_jac2(_ptr(f), int(Nx), int(Ny), float(voxel_size_x), float(voxel_size_y),
                 _ptr(out)))
    return out


def fold_mask_3D(field, Nx, Ny, Nz, voxel_size_x, voxel_size_y, voxel_size_z, e1=0.0):
    """Voxel-wise fold mask (J < e1) and its count, from the same FP64 Jacobian."""
    f = _arr(field, np.float64)
    mask = np.empty((int(Nx), int(Ny), int(Nz)), np.uint8)
    cnt = ctypes.c_longlong(0)
    _check(_fold3(_ptr(f), int(Nx), int(Ny), int(Nz), float(voxel_size_x),
                  float(voxel_size_y), float(voxel_size_z), float(e1), _ptr(mask),
                  ctypes.byref(cnt)))
    return mask.astype(bool), int(cnt.value)


_resize = _sig("binder_resize_field", _c_int, _vp, _c_int, _c_int, _c_int, _c_int, _c_int,
               _c_int, _c_int, _c_double, _vp)


def resize_field(field, output_shape, scale=1.0):
    """Field [nx,ny,nz(,comps)] resampled to output_shape (3 ints) on the GPU
    as skimage.transform.resize(order=1) does, times `scale`
    (BINDER/FastRegistration.py:566-568).  float64."""
    f = _arr(field, np.float64)
    if f.ndim not in (3, 4) or len(output_shape) != 3:
        raise ValueError("field must be [nx,ny,nz] or [nx,ny,nz,comps], shape 3 ints")
    comps = 1 if f.ndim == 3 else f.shape[3]
    shp = tuple(int(v) for v in output_shape)
    out = np.empty(shp + f.shape[3:], np.float64)
    _check(_resize(_ptr(f), f.shape[0], f.shape[1], f.shape[2], comps, *shp, float(scale),
                   _ptr(out)))
    return out


_resize_grid = _sig("binder_resize_index_grid", _c_int, _c_int, _c_int, _c_int, _c_int,
                    _c_int, _c_int, _c_int, _c_double, _vp)


def resize_index_grid(shape, comps, output_shape, scale=1.0):
    """resize_field of the index grid of `shape` (3 ints; component c of a
    voxel is its index along axis c, comps 2 or 3), made on the GPU instead
    of uploaded: [*output_shape, comps] float64."""
    shp = tuple(int(v) for v in shape)
    oshp = tuple(int(v) for v in output_shape)
    if len(shp) != 3 or len(oshp) != 3:
        raise ValueError("shape and output_shape must be 3 ints")
    out = np.empty(oshp + (int(comps),), np.float64)
    _check(_resize_grid(*shp, int(comps), *oshp, float(scale), _ptr(out)))
    return out


_minmax = _sig("binder_min_max", _c_int, _vp, ctypes.c_longlong, _P(_c_double), _P(_c_double))
_hist_uniform = _sig("binder_histogram_uniform", _c_int, _vp, ctypes.c_longlong, _c_int, _vp,
                     _vp)
_bin_floor = _sig("binder_bin_floor", _c_int, _vp, ctypes.c_longlong, _c_double, _c_double,
                  _c_double, _c_int, _vp)


def histogram_uniform(a, bins):
    """numpy.histogram(a, bins) for an integer bin count, counted over all
    host cores: (counts int64, edges float64)."""
    x = _arr(a, np.float64).ravel()
    lo, hi = _c_double(0.0), _c_double(0.0)
    _check(_minmax(_ptr(x), x.size, ctypes.byref(lo), ctypes.byref(hi)))
    if not (np.isfinite(lo.value) and np.isfinite(hi.value)):
        raise ValueError(f"autodetected range of [{lo.value}, {hi.value}] is not finite")
    first, last = lo.value, hi.value
    if first == last:
        first, last = first - 0.5, last + 0.5
    edges = np.linspace(first, last, int(bins) + 1, endpoint=True)
    counts = np.zeros(int(bins), np.int64)
    _check(_hist_uniform(_ptr(x), x.size, int(bins), _ptr(edges), _ptr(counts)))
    return counts, edges


def bin_floor(a, lo, hi, width, number_of_bins):
    """floor((a - lo) / width) as int32, 0 below lo and number_of_bins - 1
    above hi, over all host cores."""
    x = _arr(a, np.float64)
    out = np.empty(x.shape, np.int32)
    _check(_bin_floor(_ptr(x), x.size, float(lo), float(hi), float(width), int(number_of_bins),
                      _ptr(out)))
    return out


_at_set = _sig("binder_autotune_set", _c_int, _c_int)
_at_clear = _sig("binder_autotune_clear", _c_int)
_at_report = _sig("binder_autotune_report", _c_int, ctypes.c_char_p, _c_int, _P(_c_int))


def autotune(mode=1):
    """Launch-shape autotuner: 1 (or True) tunes -- the first run of a
    tunable kernel per problem size times its candidate shapes and keeps the
    fastest; 0 (False) turns it off; 2 takes each tunable's last non-default
    shape untimed (to check that other shapes give the same results).
    Results are bit for bit those of the default shapes."""
    _check(_at_set(int(mode)))


def autotune_clear():
    _check(_at_clear())


def autotune_report():
    """[(op, key, choice, default, ms_choice, ms_default)] of the tuned
    problems."""
    need = _c_int(0)
    _check(_at_report(None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(max(need.value, 1))
    _check(_at_report(buf, len(buf), ctypes.byref(need)))
    rows = []
    for line in buf.value.decode().splitlines():
        op, key, ch, df, mc, md = line.split()
        rows.append((op, int(key), int(ch), int(df), float(mc), float(md)))
    return rows


def exported_symbols():
    """Names of the C entry points this module binds (for ABI checks)."""
    return [
        "binder_last_error", "binder_backend", "binder_philox4x32_10", "binder_resize_field",
        "binder_resize_index_grid",
        "binder_topology_correction_2d", "binder_spline_weights",
        "binder_registration_correct_topology", "binder_registration_create_resampled",
        "binder_dct_prior_eigenvalues",
        "binder_registration_create",
        "binder_registration_destroy", "binder_registration_set_mi_likelihood",
        "binder_registration_set_ssd_likelihood", "binder_registration_get_ssd_parameters",
        "binder_registration_set_affine_transformation",
        "binder_registration_set_rigid_transformation",
        "binder_registration_get_linear_parameters",
        "binder_registration_set_nonlinear_transformation", "binder_registration_em",
        "binder_registration_sampler", "binder_registration_sampler_run",
        "binder_registration_sampler_result", "binder_registration_get_final_locations",
        "binder_registration_get_dream_locations", "binder_registration_set_dream_locations",
        "binder_registration_get_likelihood_parameters", "binder_registration_get_shape",
        "binder_registration_get_gamma", "binder_batch_create", "binder_batch_destroy",
        "binder_batch_em", "binder_batch_set", "binder_batch_run", "binder_min_max",
        "binder_histogram_uniform", "binder_bin_floor", "binder_autotune_set",
        "binder_autotune_clear", "binder_autotune_report",
        "binder_registration_get_objective_trace", "binder_registration_e_step",
        "binder_registration_set_profiling", "binder_registration_get_kernel_times",
        "binder_nlt_create", "binder_nlt_destroy", "binder_nlt_smooth_deformation",
        "binder_nlt_sample_deformation_field", "binder_nlt_up_sample",
        "binder_jacobian_determinant_3d", "binder_jacobian_determinant_2d",
        "binder_fold_mask_3d", "binder_topology_correction_3d",
    ]
