"""HIP DCT smoothing and DCT-domain sampling against closed forms and the
oracle (reference tests: BINDER/tests/cpu/test_smoothing_and_log_cost.py,
test_sampling_DCT_3D.py, test_sampling_DCT_3D_mean.py)."""
import numpy as np
import pytest
from scipy.fft import dctn, idctn

from binder_amd.utils import initialize_fourier_domain_smoothing

pytestmark = pytest.mark.gpu


def _nlt(core, sizes, gamma, sigma, order=2):
    D = initialize_fourier_domain_smoothing(sizes, True, order, len(sizes))
    nz = sizes[2] if len(sizes) == 3 else 1
    return core.NonLinearTransformation(gamma, gamma, gamma, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0,
                                        D.ravel(), sizes[0], sizes[1], nz, 1, sigma, 0.0), D


@pytest.mark.parametrize("sizes", [[5, 10], [15, 27], [11, 20, 7], [10, 13, 30],
                                   [64, 48, 40], [128, 96, 96], [160, 192, 160]])
def test_smooth_deformation_closed_form(core, orc, sizes):
    """Ortho DCT-II -> p = c/(1+gamma D) -> DCT-III; cost 0.5 gamma sum D p^2."""
    print(f"smooth_deformation sizes={sizes}", flush=True)
    rng = np.random.default_rng(len(sizes) * 100 + sizes[0])
    gamma = 2.0
    obj, D = _nlt(core, sizes, gamma, 1.0)
    N = len(sizes)
    nz = sizes[2] if N == 3 else 1
    field = rng.standard_normal((sizes[0], sizes[1], nz, N))
    cost, out = obj.smooth_deformation(field.ravel())
    exp_cost = 0.0
    for n in range(N):
        data = field[..., n] if N == 3 else field[:, :, 0, n]
        c = dctn(data, norm="ortho") / (1 + gamma * D)
        exp_cost += 0.5 * gamma * np.sum(D * c ** 2)
        exp = idctn(c, norm="ortho")
        got = out[..., n] if N == 3 else out[:, :, 0, n]
        assert np.max(np.abs(got - exp)) <= 2e-5 * max(1.0, np.abs(exp).max())
    assert abs(cost - exp_cost) <= 1e-5 * abs(exp_cost)
    # and the oracle agrees with the same closed form
    o, _ = _nlt(orc, sizes, gamma, 1.0)
    oc, oo = o.smooth_deformation(field.ravel())
    assert abs(oc - exp_cost) <= 1e-10 * abs(exp_cost)


def _prior_covariance(Nx, Ny, Nz, gamma, sigma, order):
    # sigma^2 (I + gamma sigma^2 Gamma^T Gamma)^-1 with Neumann first/second
    # differences (test_sampling_DCT_3D.py:24-50), x fastest (Fortran order)
    def g1(n):
        G = np.diag(np.ones(n)) - np.diag(np.ones(n - 1), -1)
        return G[1:, :]
    gp = gamma * sigma ** 2
    N = Nx * Ny * Nz
    Ix, Iy, Iz = np.eye(Nx), np.eye(Ny), np.eye(Nz)
    if order == 1:
        R = np.kron(Iz, np.kron(Iy, g1(Nx)))
        C = np.kron(Iz, np.kron(g1(Ny), Ix))
        Dd = np.kron(g1(Nz), np.kron(Ix, Iy))
        tmp = np.eye(N) + gp * (R.T @ R + C.T @ C + Dd.T @ Dd)
    else:
        G = (np.kron(Iz, np.kron(Iy, g1(Nx).T @ g1(Nx))) +
             np.kron(Iz, np.kron(g1(Ny).T @ g1(Ny), Ix)) +
             np.kron(g1(Nz).T @ g1(Nz), np.kron(Ix, Iy)))
        tmp = np.eye(N) + gp * G.T @ G
    return sigma ** 2 * np.linalg.inv(tmp)


@pytest.mark.parametrize("order", [1, 2])
def test_sample_deformation_field_covariance(core, order):
    """Empirical covariance of 1e5 DCT-domain samples equals the analytic prior
    covariance (test_sampling_DCT_3D.py:101-107 tolerance rtol=atol=1e-3)."""
    print(f"sampleDeformationField covariance order={order}", flush=True)
    Nx, Ny, Nz = 3, 5, 4
    gamma, sigma = 5.0, 0.34
    D = initialize_fourier_domain_smoothing([Nx, Ny, Nz], True, order, 3)
    obj = core.NonLinearTransformation(gamma, gamma, gamma, 1, 1, 1, 1, 1, 1, D.ravel(),
                                       Nx, Ny, Nz, 1, sigma, 0.0)
    S = 100000
    s = obj.sampleDeformationField(np.zeros((Nx, Ny, Nz, 3)).ravel(), S)
    x = s[..., 0].reshape(S, Nx * Ny * Nz, order="F").astype(np.float64)
    emp = x.T @ x / S
    cov = _prior_covariance(Nx, Ny, Nz, gamma, sigma, order)
    assert np.allclose(emp, cov, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("order", [1, 2])
def test_sampler_deformation_mean_pipeline(core, order):
    """Posterior mean (1/sigma^2) Sigma t through sampleDeformationField and
    through the full sampler with only the deformation block sampled
    (test_sampling_DCT_3D_mean.py:131-150)."""
    print(f"sampler mean pipeline order={order}", flush=True)
    rng = np.random.default_rng(12345)
    Nx, Ny, Nz = 5, 3, 4
    gamma = 5.0
    sigma = np.sqrt(9.0 / (8.0 * np.pi))
    X, Y, Z = np.meshgrid(np.linspace(-1, 1, Nx), np.linspace(-1, 1, Ny),
                          np.linspace(-1, 1, Nz), indexing="ij")
    T = np.round(((X / .5) ** 2 + (Y / .7) ** 2 + (Z / .4) ** 2 < 1) +
                 rng.standard_normal((Nx, Ny, Nz)) * 100)
    cov = _prior_covariance(Nx, Ny, Nz, gamma, sigma, order)
    mean = (cov @ T.reshape(-1, 1, order="F") / sigma ** 2).reshape((Nx, Ny, Nz), order="F")
    D = initialize_fourier_domain_smoothing([Nx, Ny, Nz], True, order, 3)
    Tt = np.zeros((Nx, Ny, Nz, 3))
    Tt[..., 0] = T
    obj = core.NonLinearTransformation(gamma, gamma, gamma, 1, 1, 1, 1, 1, 1, D.ravel(),
                                       Nx, Ny, Nz, 1, sigma, 0.0)
    S = 100000
    s = obj.sampleDeformationField(Tt.ravel(), S)
    m1 = s[..., 0].astype(np.float64).mean(0)
    assert np.allclose(mean, m1, rtol=1e-3, atol=1e-2)
    empty = np.zeros((Nx, Ny, Nz, 3))
    reg = core.Registration(final_locations=empty.ravel(), threads=1,
                            voxel_pos=empty.ravel(), node_pos=empty.ravel(), Nx_v=Nx,
                            Ny_v=Ny, Nz_v=Nz, Nx_n=Nx, Ny_n=Ny, Nz_n=Nz, spline_order=3,
                            non_zero_voxels=Nx * Ny * Nz)
    reg.setNonLinearTransformation(gamma_x=gamma, gamma_y=gamma, gamma_z=gamma,
                                   voxels_size_x=1.0, voxels_size_y=1.0, voxels_size_z=1.0,
                                   nodes_size_x=1.0, nodes_size_y=1.0, nodes_size_z=1.0,
                                   D=D.ravel())
    reg.setMILikelihood(alpha=2, theta=np.zeros(32 * 32), K=32, L=32, threads=1,
                        binned_nodes=np.zeros(Nx * Ny * Nz), binned_voxels=np.zeros(Nx * Ny * Nz))
    reg.set_dream_locations(Tt.ravel())
    m2, _ = reg.sampler(seed=12345, N_b=0, N_s=S, sample_gamma=0, sample_gamma_every=100,
                        sample_posteriors=0, sample_likelihood_parameters=0,
                        sample_transformation_parameters=1)
    assert np.allclose(mean, m2[..., 0], rtol=1e-3, atol=1e-2)


@pytest.mark.parametrize("shape_in,shape_out", [((10, 12, 8), (20, 24, 16)),
                                                ((24, 20, 1), (48, 40, 1)),
                                                ((13, 7, 9), (26, 14, 18))])
def test_up_sample_matches_oracle(core, orc, shape_in, shape_out):
    """DCT zero-pad resampling (NonLinearTransformation.cpp:366-430)."""
    print(f"up_sample {shape_in} -> {shape_out}", flush=True)
    rng = np.random.default_rng(sum(shape_in))
    N = 2 if shape_in[2] == 1 else 3
    f = rng.standard_normal(shape_in + (N,))
    g = core.NonLinearTransformation.up_sample(f.ravel(), *shape_in, *shape_out, 1)
    o = orc.NonLinearTransformation.up_sample(f.ravel(), *shape_in, *shape_out, 1)
    assert g.shape == o.shape == shape_out + (N,)
    assert np.max(np.abs(g - o)) <= 2e-5 * max(1.0, np.abs(o).max())
