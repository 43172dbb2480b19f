"""ORACLE -- test infrastructure, not product code.

ctypes wrapper of oracle/liborc.so with the same Python surface as
binder_amd._core (Registration, NonLinearTransformation,
topology_correction_3D, compute_jacobian_determinant_3D/2D) so tests can run
one scenario through the HIP backend and through the FP64 CPU oracle.
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
REF_LIB = os.path.join(HERE, "_ref", "libbinder_ref.so")


def build(ref=False):
    """Compile the oracle (and, when the reference sources exist, _ref)."""
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def load():
    if not os.path.exists(LIB):
        build()
    return ctypes.CDLL(LIB)


_lib = load()
_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int


def _sig(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_create = _sig("orc_reg_create", _vp, _vp, _i, _vp, _i, _i, _i, _i, _i, _i, _i, _d)
_destroy = _sig("orc_reg_destroy", None, _vp)
_set_mi = _sig("orc_reg_set_mi", None, _vp, _d, _vp, _i, _i, _vp, _vp)
_set_nlt = _sig("orc_reg_set_nlt", None, _vp, *([_d] * 9), _vp)
_set_ssd = _sig("orc_reg_set_ssd", None, _vp, _i, _d, _i, _d, _d, _vp, _vp, _d, _d)
_get_ssd = _sig("orc_reg_get_ssd", None, _vp, _vp)
_set_affine = _sig("orc_reg_set_affine", None, _vp, _vp, _vp)
_set_rigid = _sig("orc_reg_set_rigid", None, _vp, _vp, _vp)
_get_linear = _sig("orc_reg_get_linear", _i, _vp, _vp)
_constrain = _sig("orc_reg_constrain", _d, _vp)
_ssd_update = _sig("orc_reg_ssd_update", None, _vp, _i)
_ssd_cost = _sig("orc_reg_ssd_cost", _d, _vp)
_ssd_seed = _sig("orc_reg_ssd_seed", None, _vp, _i)
_em = _sig("orc_reg_em", _i, _vp, _i, _d, _i, _i)
_estep = _sig("orc_reg_estep", _d, _vp, _vp)
_theta_update = _sig("orc_reg_theta_update", None, _vp)
_theta_cost = _sig("orc_reg_theta_cost", _d, _vp)
_sampler = _sig("orc_reg_sampler", None, _vp, *([_i] * 8), _vp, _vp)
_get_final = _sig("orc_reg_get_final", None, _vp, _vp)
_get_dream = _sig("orc_reg_get_dream", None, _vp, _vp)
_set_dream = _sig("orc_reg_set_dream", None, _vp, _vp)
_get_theta = _sig("orc_reg_get_theta", None, _vp, _vp)
_get_trace = _sig("orc_reg_get_trace", _i, _vp, _vp, _i)
_get_gamma = _sig("orc_reg_get_gamma", None, _vp, _vp)
_nlt_create = _sig("orc_nlt_create", _vp, *([_d] * 9), _vp, _i, _i, _i, _i, _d, _d)
_nlt_destroy = _sig("orc_nlt_destroy", None, _vp)
_nlt_smooth = _sig("orc_nlt_smooth", _d, _vp, _vp, _vp)
_nlt_sample = _sig("orc_nlt_sample", None, _vp, _vp, _i, ctypes.c_longlong, _vp)
_up = _sig("orc_up_sample", None, _vp, _i, _i, _i, _i, _i, _i, _i, _vp)
_jac3 = _sig("orc_jacobian_3d", None, _vp, _i, _i, _i, _d, _d, _d, _vp)
_jac2 = _sig("orc_jacobian_2d", None, _vp, _i, _i, _d, _d, _vp)
_topo = _sig("orc_topology_3d", None, _vp, _d, _d, _i, _i, _vp, _i, _i, _i, _vp, _vp)
_corner = _sig("orc_corner_jacobians_3d", None, _vp, _i, _i, _i, _vp, _d, _d, _vp, _vp)
_tgrad = _sig("orc_topo_gradients", None, _vp, _i, _i, _i, _vp, _vp)
_tlimit = _sig("orc_topo_limit", None, _vp, _vp, _i, _i, _i, _d, _d)
_tint = _sig("orc_topo_integrate", None, _vp, _vp, _i, _i, _i, _vp, _vp)
_topo2 = _sig("orc_topology_2d", None, _vp, _d, _d, _i, _i, _vp, _i, _i, _vp, _vp)
_corner2 = _sig("orc_corner_jacobians_2d", None, _vp, _i, _i, _vp, _d, _d, _vp, _vp)
_t2grad = _sig("orc_topo2_gradients", None, _vp, _i, _i, _vp, _vp)
_t2limit = _sig("orc_topo2_limit", None, _vp, _vp, _i, _i, _d, _d)
_t2int = _sig("orc_topo2_integrate", None, _vp, _vp, _i, _i, _vp, _vp)
_spline = _sig("orc_spline", None, _i, _d, _vp)
_dct3 = _sig("orc_dct3", None, _vp, _i, _i, _i, _i, _i)


def _a(x, dt=np.float64):
    return np.ascontiguousarray(np.asarray(x), dtype=dt)


def _p(a):
    return a.ctypes.data_as(_vp)


class Registration:
    def __init__(self, final_locations, threads, voxel_pos, node_pos, Nx_v, Ny_v, Nz_v,
                 Nx_n, Ny_n, Nz_n, spline_order, non_zero_voxels):
        self._fl = _a(final_locations)
        self._vp = _a(voxel_pos)
        self._shape = (int(Nx_v), int(Ny_v), int(Nz_v), 2 if int(Nz_v) == 1 else 3)
        self._h = _create(_p(self._fl), int(threads), _p(self._vp), int(Nx_v), int(Ny_v),
                          int(Nz_v), int(Nx_n), int(Ny_n), int(Nz_n), int(spline_order),
                          float(non_zero_voxels))
        self._KL = None

    def __del__(self):
        if getattr(self, "_h", None):
            _destroy(self._h)
            self._h = None

    def setMILikelihood(self, alpha, theta, K, L, threads, binned_nodes, binned_voxels):
        self._th = _a(theta)
        self._bn = _a(binned_nodes, np.int32)
        self._bv = _a(binned_voxels, np.int32)
        _set_mi(self._h, float(alpha), _p(self._th), int(K), int(L), _p(self._bn), _p(self._bv))
        self._KL = (int(K), int(L))

    def setSSDLikelihood(self, bins, threads, sigma_sq, scale_shift, scale, shift,
                         binned_nodes, binned_voxels, alpha_0, beta_0):
        self._bn = _a(binned_nodes, np.int32)
        self._bv = _a(binned_voxels, np.int32)
        _set_ssd(self._h, int(bins), float(sigma_sq), int(scale_shift), float(scale),
                 float(shift), _p(self._bn), _p(self._bv), float(alpha_0), float(beta_0))
        self._KL = (int(bins), int(bins))
        self._ssd = True

    def get_ssd_parameters(self):
        """(scale, shift, sigma^2) of the SSD likelihood."""
        out = np.empty(3)
        _get_ssd(self._h, _p(out))
        return out.tolist()

    def setAffineTransformation(self, A, t):
        self._A = _a(A)
        self._t = _a(t)
        _set_affine(self._h, _p(self._A), _p(self._t))

    def setRigidTransformation(self, R, t):
        self._R = _a(R)
        self._t = _a(t)
        _set_rigid(self._h, _p(self._R), _p(self._t))

    def constrain(self):
        """One transformation step from the current dream locations (the
        prior cost it returns)."""
        return _constrain(self._h)

    def ssd_update(self, draw=False, seed=None):
        """One SSD parameter update (or draw) from the last e_step's histogram."""
        if seed is not None:
            _ssd_seed(self._h, int(seed))
        _ssd_update(self._h, int(bool(draw)))

    def ssd_cost(self):
        return _ssd_cost(self._h)

    def get_linear_parameters(self):
        """The fitted affine ((N+1)^2, row-major) or rigid (N^2) matrix, then t."""
        out = np.empty(32)
        n = _get_linear(self._h, _p(out))
        return out[:n].tolist()

    def setNonLinearTransformation(self, gamma_x, gamma_y, gamma_z, voxels_size_x,
                                   voxels_size_y, voxels_size_z, nodes_size_x, nodes_size_y,
                                   nodes_size_z, D):
        self._D = _a(D)
        _set_nlt(self._h, gamma_x, gamma_y, gamma_z, voxels_size_x, voxels_size_y,
                 voxels_size_z, nodes_size_x, nodes_size_y, nodes_size_z, _p(self._D))

    def EM(self, max_EM_iterations, convergence_th, update_likelihood_parameters, debug):
        return _em(self._h, int(max_EM_iterations), float(convergence_th),
                   int(bool(update_likelihood_parameters)), int(bool(debug)))

    def sampler(self, seed, N_b, N_s, sample_gamma, sample_gamma_every, sample_posteriors,
                sample_likelihood_parameters, sample_transformation_parameters):
        Nx, Ny, Nz, N = self._shape
        mean = np.empty((Nx, Ny, Nz, N))
        cov = np.empty((Nx, Ny, Nz, N, N))
        _sampler(self._h, int(seed), int(N_b), int(N_s), int(sample_gamma),
                 int(sample_gamma_every), int(sample_posteriors),
                 int(sample_likelihood_parameters), int(sample_transformation_parameters),
                 _p(mean), _p(cov))
        return mean, cov

    def get_final_locations_as_numpy_array(self):
        out = np.empty(self._shape)
        _get_final(self._h, _p(out))
        return out

    def get_dream_locations(self):
        out = np.empty(self._shape)
        _get_dream(self._h, _p(out))
        return out.ravel().tolist()

    def set_dream_locations(self, dream_locations):
        d = _a(dream_locations)
        _set_dream(self._h, _p(d))

    def get_likelihood_parameters(self):
        if getattr(self, "_ssd", False):
            return self.get_ssd_parameters()[:2]  # [scale, shift]
        out = np.empty(self._KL[0] * self._KL[1])
        _get_theta(self._h, _p(out))
        return out.tolist()

    def get_shape(self):
        return list(self._shape)

    def get_gamma(self):
        out = np.empty(3)
        _get_gamma(self._h, _p(out))
        return out.tolist()

    def get_objective_trace(self):
        rows = _get_trace(self._h, None, 0)
        out = np.empty((max(rows, 1), 4))
        _get_trace(self._h, _p(out), rows)
        return out[:rows]

    def e_step(self, return_histogram=False):
        H = np.empty(self._KL)
        mll = _estep(self._h, _p(H))
        return (mll, H) if return_histogram else mll

    def theta_update(self):
        _theta_update(self._h)

    def theta_cost(self):
        return _theta_cost(self._h)


class NonLinearTransformation:
    def __init__(self, gamma_x, gamma_y, gamma_z, voxels_size_x, voxels_size_y, voxels_size_z,
                 nodes_size_x, nodes_size_y, nodes_size_z, D, Nx_v, Ny_v, Nz_v, threads,
                 sigma_spline, spline_offset):
        self._D = _a(D)
        self._shape = (int(Nx_v), int(Ny_v), int(Nz_v), 2 if int(Nz_v) == 1 else 3)
        self._h = _nlt_create(gamma_x, gamma_y, gamma_z, voxels_size_x, voxels_size_y,
                              voxels_size_z, nodes_size_x, nodes_size_y, nodes_size_z,
                              _p(self._D), int(Nx_v), int(Ny_v), int(Nz_v), int(threads),
                              float(sigma_spline), float(spline_offset))

    def __del__(self):
        if getattr(self, "_h", None):
            _nlt_destroy(self._h)
            self._h = None

    def get_shape(self):
        return list(self._shape)

    def smooth_deformation(self, deformation):
        d = _a(deformation)
        out = np.empty(self._shape)
        c = _nlt_smooth(self._h, _p(d), _p(out))
        return c, out

    def sampleDeformationField(self, deformationField, number_of_samples, seed=-1):
        d = _a(deformationField)
        out = np.empty((int(number_of_samples),) + self._shape)
        _nlt_sample(self._h, _p(d), int(number_of_samples), int(seed), _p(out))
        return out

    @staticmethod
    def up_sample(deformation_field, N_x, N_y, N_z, N_x_o, N_y_o, N_z_o, threads):
        f = _a(deformation_field)
        N = 2 if int(N_z_o) == 1 else 3
        out = np.empty((int(N_x_o), int(N_y_o), int(N_z_o), N))
        _up(_p(f), int(N_x), int(N_y), int(N_z), int(N_x_o), int(N_y_o), int(N_z_o),
            int(threads), _p(out))
        return out


class topology_correction_3D:
    def __init__(self):
        self.last_stats = None

    def correct_topology(self, h_init, e1, e2, max_outer_it, max_it, threads, voxel_sizes,
                         Nx_v, Ny_v, Nz_v):
        h = _a(h_init)
        vs = _a(voxel_sizes)
        out = np.empty((int(Nx_v), int(Ny_v), int(Nz_v), 3))
        st = np.zeros(6)
        _topo(_p(h), float(e1), float(e2), int(max_outer_it), int(max_it), _p(vs), int(Nx_v),
              int(Ny_v), int(Nz_v), _p(out), _p(st))
        self.last_stats = st
        return out


class topology_correction_2D:
    def __init__(self):
        self.last_stats = None

    def correct_topology(self, h_init, e1, e2, max_outer_it, max_it, threads, voxel_sizes,
                         Nx_v, Ny_v):
        h = _a(h_init)
        vs = _a(voxel_sizes)
        out = np.empty((int(Nx_v), int(Ny_v), 2))
        st = np.zeros(6)
        _topo2(_p(h), float(e1), float(e2), int(max_outer_it), int(max_it), _p(vs), int(Nx_v),
               int(Ny_v), _p(out), _p(st))
        self.last_stats = st
        return out


def corner_jacobians_2D(field, Nx, Ny, voxel_sizes, e1, e2):
    f = _a(field)
    Jc = np.empty((Nx - 1) * (Ny - 1) * 4)
    cnt = np.zeros(2, np.int64)
    _corner2(_p(f), Nx, Ny, _p(_a(voxel_sizes)), float(e1), float(e2), _p(Jc), _p(cnt))
    return Jc, cnt


def topo2_gradients(w_phys, shape, voxel_sizes):
    g = np.empty(tuple(shape) + (2, 2))
    _t2grad(_p(_a(w_phys)), *map(int, shape), _p(_a(voxel_sizes)), _p(g))
    return g


def topo2_limit(g, Jc, shape, e1, e2):
    out = _a(g).copy()
    _t2limit(_p(out), _p(_a(Jc)), *map(int, shape), float(e1), float(e2))
    return out


def topo2_integrate(g, mean2, shape, voxel_sizes):
    out = np.empty(tuple(shape) + (2,))
    _t2int(_p(_a(g)), _p(_a(mean2)), *map(int, shape), _p(_a(voxel_sizes)), _p(out))
    return out


def compute_jacobian_determinant_3D(field, Nx, Ny, Nz, voxel_size_x, voxel_size_y,
                                    voxel_size_z):
    f = _a(field)
    out = np.empty((int(Nx), int(Ny), int(Nz)))
    _jac3(_p(f), int(Nx), int(Ny), int(Nz), voxel_size_x, voxel_size_y, voxel_size_z, _p(out))
    return out


def compute_jacobian_determinant_2D(field, Nx, Ny, voxel_size_x, voxel_size_y):
    f = _a(field)
    out = np.empty((int(Nx), int(Ny)))
    _jac2(_p(f), int(Nx), int(Ny), voxel_size_x, voxel_size_y, _p(out))
    return out


def corner_jacobians_3D(field, Nx, Ny, Nz, voxel_sizes, e1, e2):
    f = _a(field)
    vs = _a(voxel_sizes)
    Jc = np.empty((Nx - 1) * (Ny - 1) * (Nz - 1) * 8)
    cnt = np.zeros(2, np.int64)
    _corner(_p(f), Nx, Ny, Nz, _p(vs), float(e1), float(e2), _p(Jc), _p(cnt))
    return Jc, cnt


def topo_gradients(w_phys, shape, voxel_sizes):
    """Periodic forward differences [Nx,Ny,Nz,3(component),3(axis)] / vs[component]."""
    w = _a(w_phys)
    g = np.empty(tuple(shape) + (3, 3))
    _tgrad(_p(w), *map(int, shape), _p(_a(voxel_sizes)), _p(g))
    return g


def topo_limit(g, Jc, shape, e1, e2):
    """Sequential in-place Jacobian limiting of gradient field g (copy returned)."""
    out = _a(g).copy()
    _tlimit(_p(out), _p(_a(Jc)), *map(int, shape), float(e1), float(e2))
    return out


def topo_integrate(g, mean3, shape, voxel_sizes):
    """Periodic least-squares integration of g, scaled by vs, plus mean3."""
    out = np.empty(tuple(shape) + (3,))
    _tint(_p(_a(g)), _p(_a(mean3)), *map(int, shape), _p(_a(voxel_sizes)), _p(out))
    return out


def resize_field(field, output_shape, scale=1.0):
    """Host restatement of the pyramid resampling (same surface as _core)."""
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from binder_amd import resample
    return resample.resize(field, output_shape, order=1) * scale


def spline(order, d):
    out = np.empty(4)
    _spline(int(order), float(d), _p(out))
    return out


def dct3(x, inverse=False, threads=1):
    y = _a(x).copy()
    shp = y.shape + (1,) * (3 - y.ndim)
    _dct3(_p(y), shp[0], shp[1], shp[2], int(inverse), int(threads))
    return y
