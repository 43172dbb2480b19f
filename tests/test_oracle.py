"""The FP64 CPU oracle pinned against the reference's own arithmetic (golden
vectors written by tests/golden/make_golden.py from the reference sources),
against the closed forms of the reference's tests, and against brute force.

Reference tests restated here: BINDER/tests/cpu/test_smoothing_and_log_cost.py,
test_DCT_smoothing_3D.py, test_cost_DCT.py, test_check_spline.py,
test_sampling_DCT_3D.py, test_sampling_DCT_3D_mean.py.
"""
import ctypes
import os

import numpy as np
import pytest
from numpy.fft import fftn
from scipy.fft import dct, dctn, idctn
from scipy.stats import dirichlet

from binder_amd.utils import initialize_fourier_domain_smoothing

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))


# ---------------------------------------------------------------------------
# golden vectors from the reference's compiled utils.cpp / Spline.cpp /
# MILikelihood.cpp
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("i", range(4))
def test_jacobian_3d_golden_bitexact(orc, i):
    f = GOLD[f"jac3_{i}_field"]
    vs = GOLD[f"jac3_{i}_vs"]
    J = orc.compute_jacobian_determinant_3D(f.ravel(), *f.shape[:3], *vs)
    assert np.array_equal(J, GOLD[f"jac3_{i}_J"])


@pytest.mark.parametrize("i", range(3))
def test_jacobian_2d_golden_bitexact(orc, i):
    f = GOLD[f"jac2_{i}_field"]
    vs = GOLD[f"jac2_{i}_vs"]
    J = orc.compute_jacobian_determinant_2D(f.ravel(), *f.shape[:2], *vs)
    assert np.array_equal(J, GOLD[f"jac2_{i}_J"])


def test_spline_golden(orc):
    ds = GOLD["spline_d"]
    for oi, order in enumerate((1, 2, 3)):
        for j, d in enumerate(ds):
            assert np.array_equal(orc.spline(order, d), GOLD["spline_w"][oi, j])


@pytest.mark.parametrize("update", [0, 1])
def test_mi_likelihood_golden(orc, update):
    """Lookups (out-of-range nodes -> moving bin 0), accumulate_info,
    update_parameters and compute_cost through the oracle's code paths."""
    lib = orc._lib
    f = lib.orc_mi_step
    f.restype = ctypes.c_double
    K, L = int(GOLD["mi_K"]), int(GOLD["mi_L"])
    th = np.ascontiguousarray(GOLD["mi_theta_in"].ravel().copy())
    bn = np.ascontiguousarray(GOLD["mi_bn"], np.int32)
    bv = np.ascontiguousarray(GOLD["mi_bv"], np.int32)
    qv = np.ascontiguousarray(GOLD["mi_qv"], np.int32)
    qn = np.ascontiguousarray(GOLD["mi_qn"], np.int32)
    qp = np.ascontiguousarray(GOLD["mi_qp"])
    lik = np.empty(len(qv))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    cost = f(ctypes.c_double(float(GOLD["mi_alpha"])), p(th), K, L, p(bn), p(bv),
             *bn.shape, *bv.shape, len(qv), p(qv), p(qn), p(qp), p(lik), update)
    assert np.array_equal(lik, GOLD[f"mi_lik_{update}"])
    assert np.allclose(th, GOLD[f"mi_theta_out_{update}"], rtol=1e-14, atol=0)
    assert abs(cost - float(GOLD[f"mi_cost_{update}"])) <= 1e-12 * abs(cost)
    # closed form: MAP (H + alpha - 1) / row sum and -sum log Dir(theta_k; alpha)
    H = np.zeros((K, L))
    for v, n, w in zip(qv, qn, qp):
        inside = all(0 <= n[a] < bn.shape[a] for a in range(3))
        k = bn[tuple(n)] if inside else 0
        H[k, bv.ravel()[v]] += w
    alpha = float(GOLD["mi_alpha"])
    exp = (H + alpha - 1) / (H + alpha - 1).sum(1, keepdims=True) if update else \
        GOLD["mi_theta_in"]
    assert np.allclose(th.reshape(K, L), exp, rtol=1e-13)
    exp_cost = -sum(dirichlet.logpdf(exp[k], np.full(L, alpha)) for k in range(K))
    assert abs(cost - exp_cost) <= 1e-9 * abs(exp_cost)


def test_ref_build_bitexact_random(orc):
    """Where oracle/_ref was built from the reference sources, the oracle's
    Jacobian equals the reference's on fresh random fields."""
    if not os.path.exists(orc.REF_LIB):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    ref = ctypes.CDLL(orc.REF_LIB)
    rng = np.random.default_rng(99)
    for shape in [(8, 9, 10), (3, 17, 5)]:
        f = np.ascontiguousarray(np.stack(np.indices(shape, dtype=np.float64), -1) +
                                 rng.normal(0, 0.9, shape + (3,)))
        a = orc.compute_jacobian_determinant_3D(f.ravel(), *shape, 1.3, 0.7, 1.1)
        b = np.empty(shape)
        p = lambda x: x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        d = ctypes.c_double
        ref.ref_jacobian_3d(p(f), *shape, d(1.3), d(0.7), d(1.1), p(b))
        assert np.array_equal(a, b)


# ---------------------------------------------------------------------------
# spline invariants (test_check_spline.py:40-76)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("order", [1, 2, 3])
def test_spline_partition_of_unity(orc, order):
    for d3 in ([0.2, 0.4, 0.1], [0.0, 1.0, -0.3]):
        w = [orc.spline(order, d)[:order + 1] for d in d3]
        assert np.isclose(np.sum(w[0][:, None] * w[1][None, :]), 1.0)
        assert np.isclose(np.sum(w[0][:, None, None] * w[1][None, :, None] *
                                 w[2][None, None, :]), 1.0)
    # cubic: first moment reproduces the offset (linear precision)
    if order == 3:
        for d in (0.0, 0.3, 0.77):
            w = orc.spline(3, d)
            assert np.isclose(np.dot(w, [-1, 0, 1, 2]), d)


# ---------------------------------------------------------------------------
# orthonormal DCT and smoothing (test_smoothing_and_log_cost.py:25-146)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(5, 10, 1), (15, 27, 1), (11, 20, 7), (10, 13, 30),
                                   (160, 1, 1), (7, 11, 13)])
def test_oracle_dct_matches_scipy(orc, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    axes = tuple(i for i in range(3) if shape[i] > 1)
    c = orc.dct3(x)
    assert np.allclose(c, dctn(x, norm="ortho", axes=axes), atol=1e-12)
    assert np.allclose(orc.dct3(c, inverse=True), x, atol=1e-12)
    # one line against the dense O(N^2) orthonormal DCT-II matrix
    n = shape[0]
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    C = np.sqrt(2.0 / n) * np.cos(np.pi * (2 * j + 1) * k / (2 * n))
    C[0] /= np.sqrt(2.0)
    line = x[:, 0, 0].copy()
    assert np.allclose(orc.dct3(line.reshape(n, 1, 1)).ravel(), C @ line, atol=1e-12)
    assert np.allclose(dct(line, norm="ortho"), C @ line, atol=1e-12)


@pytest.mark.parametrize("sizes", [[5, 10], [15, 27], [11, 20, 7], [10, 13, 30]])
def test_oracle_smoothing_and_log_cost(orc, sizes):
    """Smoothed field equals scipy DCT smoothing and the mirrored-DFT smoothing;
    the prior cost equals the DFT-domain cost on the mirrored signal."""
    gamma = 2.0
    rng = np.random.default_rng(7 + sum(sizes))
    nd = len(sizes)
    nz = sizes[2] if nd == 3 else 1
    field = rng.standard_normal((sizes[0], sizes[1], nz, nd))
    Dbar, Dfull = initialize_fourier_domain_smoothing(sizes, DCT=True, finite_difference_order=2,
                                                      number_of_dimensions=nd,
                                                      return_full_matrix=True)
    obj = orc.NonLinearTransformation(gamma, gamma, gamma, 1, 1, 1, 1, 1, 1, Dbar.ravel(),
                                      sizes[0], sizes[1], nz, 1, 1.0, 0.0)
    cost, out = obj.smooth_deformation(field.ravel())
    cost_fft = 0.0
    for n in range(nd):
        data = field[..., n] if nd == 3 else field[:, :, 0, n]
        exp = idctn(dctn(data, norm="ortho") / (1 + gamma * Dbar), norm="ortho")
        got = out[..., n] if nd == 3 else out[:, :, 0, n]
        assert np.allclose(got, exp, atol=1e-12)
        m = np.vstack((data, np.flip(data, axis=0)))
        m = np.hstack((m, np.flip(m, axis=1)))
        if nd == 3:
            m = np.concatenate((m, np.flip(m, axis=2)), axis=-1)
        Cf = fftn(m) / m.size / (1 + gamma * Dfull)
        cost_fft += 0.5 * gamma * m.size * np.sum(Dfull * np.abs(Cf) ** 2) / 2 ** nd
        t = np.fft.ifftn(Cf).real * m.size
        assert np.allclose(got, t[tuple(slice(0, s) for s in data.shape)], atol=1e-10)
    assert np.isclose(cost, cost_fft, rtol=1e-10)


def _neumann_diff(n):
    G = np.diag(np.ones(n)) - np.diag(np.ones(n - 1), -1)
    return G[1:, :]


def test_prior_precision_decomposition():
    """Gamma^T Gamma = C diag(Delta) C^-1 for first and second differences
    (test_DCT_smoothing_3D.py:70-86): pins the eigenvalue array D."""
    Nx, Ny, Nz = 5, 6, 4
    C = []
    for n in (Nx, Ny, Nz):
        k = np.arange(n)
        Cn = np.cos(np.pi / n * (k[:, None] + .5) @ k[None, :])
        Cn[:, 0] /= np.sqrt(2)
        C.append(Cn)
    Ck = np.kron(C[2], np.kron(C[1], C[0]))
    D1 = initialize_fourier_domain_smoothing([Nx, Ny, Nz], True, 1, 3)
    D2 = initialize_fourier_domain_smoothing([Nx, Ny, Nz], True, 2, 3)
    G = [_neumann_diff(n) for n in (Nx, Ny, Nz)]
    I = [np.eye(n) for n in (Nx, Ny, Nz)]
    R = np.kron(I[2], np.kron(I[1], G[0]))
    Cc = np.kron(I[2], np.kron(G[1], I[0]))
    Dd = np.kron(G[2], np.kron(I[0], I[1]))
    lap1 = R.T @ R + Cc.T @ Cc + Dd.T @ Dd
    assert np.allclose(lap1, Ck @ np.diag(D1.ravel(order="F")) @ np.linalg.inv(Ck))
    G2 = (np.kron(I[2], np.kron(I[1], G[0].T @ G[0])) + np.kron(I[2], np.kron(G[1].T @ G[1], I[0])) +
          np.kron(G[2].T @ G[2], np.kron(I[0], I[1])))
    assert np.allclose(G2.T @ G2, Ck @ np.diag(D2.ravel(order="F")) @ np.linalg.inv(Ck))


@pytest.mark.parametrize("order", [1, 2])
def test_oracle_cost_equals_spatial_finite_differences(orc, order):
    """0.5 gamma sum D p^2 equals 0.5 gamma ||Gamma t_hat||^2 of the smoothed
    field in the spatial domain (test_cost_DCT.py)."""
    Nx, Ny, Nz = 9, 8, 7
    gamma = 3.0
    rng = np.random.default_rng(order)
    D = initialize_fourier_domain_smoothing([Nx, Ny, Nz], True, order, 3)
    obj = orc.NonLinearTransformation(gamma, gamma, gamma, 1, 1, 1, 1, 1, 1, D.ravel(),
                                      Nx, Ny, Nz, 1, 1.0, 0.0)
    field = rng.standard_normal((Nx, Ny, Nz, 3))
    cost, out = obj.smooth_deformation(field.ravel())
    G = [_neumann_diff(n) for n in (Nx, Ny, Nz)]
    spatial = 0.0
    for n in range(3):
        t = out[..., n]
        dx = np.einsum("ij,jkl->ikl", G[0], t)
        dy = np.einsum("ij,kjl->kil", G[1], t)
        dz = np.einsum("ij,klj->kli", G[2], t)
        if order == 1:
            spatial += np.sum(dx ** 2) + np.sum(dy ** 2) + np.sum(dz ** 2)
        else:
            lap = (np.einsum("ij,jkl->ikl", G[0].T @ G[0], t) +
                   np.einsum("ij,kjl->kil", G[1].T @ G[1], t) +
                   np.einsum("ij,klj->kli", G[2].T @ G[2], t))
            spatial += np.sum(lap ** 2)
    assert np.isclose(cost, 0.5 * gamma * spatial, rtol=1e-9)


# ---------------------------------------------------------------------------
# E-step against a direct NumPy loop written from the model
# ---------------------------------------------------------------------------
def _bspline3(t):
    return np.array([(1 - t) ** 3 / 6, 2 / 3 - t ** 2 + t ** 3 / 2,
                     2 / 3 - (1 - t) ** 2 + (1 - t) ** 3 / 2, t ** 3 / 6])


def _estep_bruteforce(d, bn, bv, theta):
    shape = bv.shape
    nd = d.shape[-1]
    K, L = theta.shape
    H = np.zeros((K, L))
    dream = np.zeros_like(d)
    mll = 0.0
    for idx in np.ndindex(*shape):
        p = d[idx]
        base = np.floor(p).astype(int)
        w = [_bspline3(p[a] - base[a]) for a in range(nd)]
        nodes, post = [], []
        for off in np.ndindex(*(4,) * nd):
            node = base + np.array(off) - 1
            inside = all(0 <= node[a] < bn.shape[a] for a in range(nd))
            k = bn[tuple(node) + (0,) * (bn.ndim - nd)] if inside else 0
            pw = np.prod([w[a][off[a]] for a in range(nd)]) * theta[k, bv[idx]]
            nodes.append((node, k))
            post.append(pw)
        post = np.array(post)
        Z = post.sum()
        mll -= np.log(Z)
        for (node, k), pw in zip(nodes, post / Z):
            H[k, bv[idx]] += pw
            dream[idx] += pw * node
    return mll, H, dream


@pytest.mark.parametrize("shape,K,L", [((5, 4, 6), 4, 4), ((7, 6, 1), 3, 5)])
def test_oracle_estep_bruteforce(orc, shape, K, L):
    from binder_amd import synth
    rng = np.random.default_rng(3)
    nd = 2 if shape[2] == 1 else 3
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)[..., :nd]
    d = pos + rng.normal(0, 1.2, pos.shape)
    d[0] += 7.5  # some voxels far outside the moving grid
    bn = rng.integers(0, K, shape).astype(np.int32)
    bv = rng.integers(0, L, shape).astype(np.int32)
    theta = rng.random((K, L)) + 0.05
    theta /= theta.sum(1, keepdims=True)
    D = initialize_fourier_domain_smoothing(list(shape[:nd]), True, 2, nd).reshape(shape)
    prob = synth.Problem(shape=shape, final_locations=d, voxel_pos=pos, binned_nodes=bn,
                         binned_voxels=bv, theta=theta, D=D, gammas=(1.0, 1.0, 1.0), K=K, L=L,
                         alpha=2.0, nnz=int(np.prod(shape)))
    r = synth.build_registration(orc, prob)
    mll, H = r.e_step(return_histogram=True)
    dream = np.asarray(r.get_dream_locations()).reshape(d.shape)
    bmll, bH, bdream = _estep_bruteforce(d[:, :, 0, :] if nd == 2 else d,
                                         bn[..., 0] if nd == 2 else bn,
                                         bv[..., 0] if nd == 2 else bv, theta)
    bdream = bdream.reshape(d.shape)
    assert np.isclose(mll, bmll, rtol=1e-12)
    assert np.allclose(H, bH, atol=1e-12)
    assert np.allclose(dream, bdream, atol=1e-12)
    assert np.isclose(H.sum(), np.prod(shape))


def test_oracle_estep_uniform_theta_reproduces_positions(orc):
    """theta = 1/L: the posterior is the spline prior, whose mean is d."""
    from binder_amd import synth
    pair = synth.make_pair((14, 12, 10), 3, 1.0, seed=2, semi=(2.0, 5.0))
    prob = synth.problem_from_pair(pair, gamma=1e-3)
    rng = np.random.default_rng(5)
    prob.final_locations = prob.voxel_pos + rng.normal(0, 0.7, prob.voxel_pos.shape)
    r = synth.build_registration(orc, prob)
    mll, H = r.e_step(return_histogram=True)
    dream = np.asarray(r.get_dream_locations()).reshape(prob.final_locations.shape)
    assert np.allclose(dream, prob.final_locations, atol=1e-12)
    assert np.isclose(mll, np.prod(prob.shape) * np.log(prob.L))


# ---------------------------------------------------------------------------
# DCT-domain sampling (test_sampling_DCT_3D.py / _mean.py) on the oracle
# ---------------------------------------------------------------------------
def _prior_cov(Nx, Ny, Nz, gamma, sigma, order):
    g = [_neumann_diff(n) for n in (Nx, Ny, Nz)]
    I = [np.eye(n) for n in (Nx, Ny, Nz)]
    N = Nx * Ny * Nz
    gp = gamma * sigma ** 2
    if order == 1:
        R = np.kron(I[2], np.kron(I[1], g[0]))
        C = np.kron(I[2], np.kron(g[1], I[0]))
        Dd = np.kron(g[2], np.kron(I[0], I[1]))
        tmp = np.eye(N) + gp * (R.T @ R + C.T @ C + Dd.T @ Dd)
    else:
        G = (np.kron(I[2], np.kron(I[1], g[0].T @ g[0])) +
             np.kron(I[2], np.kron(g[1].T @ g[1], I[0])) +
             np.kron(g[2].T @ g[2], np.kron(I[0], I[1])))
        tmp = np.eye(N) + gp * G.T @ G
    return sigma ** 2 * np.linalg.inv(tmp)


def test_oracle_sampling_covariance(orc):
    Nx, Ny, Nz, gamma, sigma = 3, 5, 4, 5.0, 0.34
    D = initialize_fourier_domain_smoothing([Nx, Ny, Nz], True, 2, 3)
    obj = orc.NonLinearTransformation(gamma, gamma, gamma, 1, 1, 1, 1, 1, 1, D.ravel(),
                                      Nx, Ny, Nz, 1, sigma, 0.0)
    S = 100000
    s = obj.sampleDeformationField(np.zeros((Nx, Ny, Nz, 3)).ravel(), S, seed=12345)
    x = s[..., 0].reshape(S, -1, order="F")
    assert np.allclose(x.T @ x / S, _prior_cov(Nx, Ny, Nz, gamma, sigma, 2), rtol=1e-3,
                       atol=1e-3)


def test_oracle_sampler_mean_pipeline(orc):
    Nx, Ny, Nz, gamma = 5, 3, 4, 5.0
    sigma = np.sqrt(9.0 / (8.0 * np.pi))
    rng = np.random.default_rng(12345)
    T = np.round(rng.standard_normal((Nx, Ny, Nz)) * 100)
    cov = _prior_cov(Nx, Ny, Nz, gamma, sigma, 1)
    mean = (cov @ T.reshape(-1, 1, order="F") / sigma ** 2).reshape((Nx, Ny, Nz), order="F")
    D = initialize_fourier_domain_smoothing([Nx, Ny, Nz], True, 1, 3)
    Tt = np.zeros((Nx, Ny, Nz, 3))
    Tt[..., 0] = T
    empty = np.zeros((Nx, Ny, Nz, 3))
    reg = orc.Registration(empty.ravel(), 1, empty.ravel(), empty.ravel(), Nx, Ny, Nz,
                           Nx, Ny, Nz, 3, Nx * Ny * Nz)
    reg.setNonLinearTransformation(gamma, gamma, gamma, 1, 1, 1, 1, 1, 1, D.ravel())
    reg.setMILikelihood(2, np.zeros(32 * 32), 32, 32, 1, np.zeros(Nx * Ny * Nz),
                        np.zeros(Nx * Ny * Nz))
    reg.set_dream_locations(Tt.ravel())
    m, _ = reg.sampler(12345, 0, 100000, 0, 100, 0, 0, 1)
    assert np.allclose(mean, m[..., 0], rtol=1e-3, atol=1e-2)


# ---------------------------------------------------------------------------
# topology correction invariants (topology_correction_3D.cpp:17-143)
# ---------------------------------------------------------------------------
def test_oracle_topology_invariants(orc):
    from binder_amd import synth
    shape = (16, 14, 12)
    rng = np.random.default_rng(58)
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    f = pos + synth.dct_displacement(shape, 3.0, rng, ncoef=4)
    t = orc.topology_correction_3D()
    out = t.correct_topology(f.ravel(), 0.01, 100, 10, 10, 1, np.ones(3), *shape)
    Jc, cnt = orc.corner_jacobians_3D(out.ravel(), *shape, np.ones(3), 0.01, 100)
    assert cnt[0] == 0 and cnt[1] == 0
    assert Jc.min() >= 0.01
    assert np.allclose(out.reshape(-1, 3).mean(0), f.reshape(-1, 3).mean(0), atol=1e-10)
    same = t.correct_topology(pos.ravel(), 0.01, 100, 10, 10, 1, np.ones(3), *shape)
    assert np.array_equal(same, pos) and t.last_stats[0] == 0


# ---------------------------------------------------------------------------
# topology correction inner steps (topology_correction_3D.cpp:172-449)
# ---------------------------------------------------------------------------
def test_topology_integrate_matches_fft_closed_form(orc):
    """Periodic least-squares integration = numpy FFT of
    sum_a conj(e^{iw_a}-1) g_a / sum_a (2 - 2 cos w_a), DC zero, x vs + mean."""
    rng = np.random.default_rng(71)
    shape = (6, 5, 7)
    vs = np.array([1.3, 0.8, 1.1])
    g = rng.standard_normal(shape + (3, 3))
    m = np.array([3.0, -2.0, 0.5])
    got = orc.topo_integrate(g, m, shape, vs)
    w = [np.exp(2j * np.pi * np.fft.fftfreq(n)) for n in shape]
    E = [w[0][:, None, None], w[1][None, :, None], w[2][None, None, :]]
    den = sum(2 - 2 * np.real(e) for e in E)
    den[0, 0, 0] = 1.0
    for n in range(3):
        U = sum(np.conj(E[a] - 1) * np.fft.fftn(g[..., n, a]) for a in range(3)) / den
        U[0, 0, 0] = 0.0
        exp = np.real(np.fft.ifftn(U)) * vs[n] + m[n]
        assert np.allclose(got[..., n], exp, atol=1e-12)


def test_topology_gradients_then_integrate_roundtrip(orc):
    """Exact periodic gradients integrate back to the field (mean restored);
    affine fields have constant forward differences A[n, a] / vs[n]."""
    rng = np.random.default_rng(72)
    shape = (7, 6, 5)
    vs = np.array([0.9, 1.2, 1.05])
    w = rng.standard_normal(shape + (3,)) * 4
    g = orc.topo_gradients(w, shape, vs)
    back = orc.topo_integrate(g, w.reshape(-1, 3).mean(0), shape, vs)
    assert np.allclose(back, w, atol=1e-10)
    A = rng.standard_normal((3, 3))
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    g = orc.topo_gradients(pos @ A.T, shape, vs)
    inner = g[:-1, :-1, :-1]
    assert np.allclose(inner, (A / vs[:, None])[None, None, None], atol=1e-12)


def _corner_columns(cell, c_vertex):
    """Voxels holding the three edge differences of the corner at vertex
    cell + c_vertex: the edge along axis a starts at the cell's lower face."""
    locs = []
    for a in range(3):
        v = [cell[k] + c_vertex[k] for k in range(3)]
        v[a] = cell[a]
        locs.append(tuple(v))
    return locs


@pytest.mark.parametrize("fold", [False, True])
def test_topology_limit_single_corner(orc, fold):
    """Limiting one flagged corner rewrites exactly its three edge-difference
    columns, whose determinant is that corner's Jacobian, as (1-a)J + aI with
    a one 0.1 step past the first blend whose determinant is in [e1, e2]."""
    rng = np.random.default_rng(73)
    shape = (5, 4, 6)
    vs = np.ones(3)
    e1, e2 = 0.01, 100.0
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    w = pos + rng.normal(0, 0.05, pos.shape)
    cell = (2, 1, 3)
    if fold:
        w[3, 2, 4, 0] -= 1.6  # pull a vertex of the cell back past its neighbour
    Jc, _ = orc.corner_jacobians_3D(w.ravel(), *shape, vs, e1, e2)
    Jc = Jc.reshape(shape[0] - 1, shape[1] - 1, shape[2] - 1, 8)
    g0 = orc.topo_gradients(w, shape, vs)
    vertices = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 0, 1), (1, 1, 0),
                (1, 1, 1)]
    corners = [int(np.argmin(Jc[cell]))] if fold else range(8)
    for c in corners:
        jc = np.full_like(Jc, 1.0)
        jc[cell + (c,)] = Jc[cell + (c,)] if fold else -1.0
        g = orc.topo_limit(g0, jc.ravel(), shape, e1, e2)
        locs = _corner_columns(cell, vertices[c])
        M = np.stack([g0[locs[a]][:, a] for a in range(3)], 1)
        assert np.isclose(np.linalg.det(M), Jc[cell + (c,)], rtol=1e-12, atol=1e-14)
        expect = g0.copy()
        alpha = 0.0
        det = np.linalg.det(M)
        while det < e1 or det > e2:
            alpha = min(alpha + 0.1, 1.0)
            det = np.linalg.det((1 - alpha) * M + alpha * np.eye(3))
        alpha = min(alpha + 0.1, 1.0)
        new = (1 - alpha) * M + alpha * np.eye(3)
        for a in range(3):
            expect[locs[a]][:, a] = new[:, a]
        assert np.allclose(g, expect, atol=1e-14, rtol=0)
        if fold:
            assert alpha > 0.15 and e1 <= np.linalg.det(new) <= e2


@pytest.mark.parametrize("shape_in,shape_out", [((6, 5, 4), (12, 10, 8)),
                                                ((7, 9, 1), (14, 18, 1)),
                                                ((5, 6, 7), (9, 11, 13))])
def test_oracle_up_sample_closed_form(orc, shape_in, shape_out):
    """NonLinearTransformation::up_sample (NonLinearTransformation.cpp:366-430):
    ortho DCT-II of each component at the input size, coefficients copied
    into the low corner of a zero array of the output size, ortho DCT-III."""
    from scipy.fft import dctn, idctn
    rng = np.random.default_rng(sum(shape_out))
    N = 2 if shape_out[2] == 1 else 3
    f = rng.standard_normal(shape_in + (N,))
    got = orc.NonLinearTransformation.up_sample(f.ravel(), *shape_in, *shape_out, 1)
    for n in range(N):
        c = np.zeros(shape_out)
        c[:shape_in[0], :shape_in[1], :shape_in[2]] = dctn(f[..., n], norm="ortho")
        assert np.allclose(got[..., n], idctn(c, norm="ortho"), atol=1e-12)


# ---------------------------------------------------------------------------
# 2D topology correction (topology_correction_2D.cpp:17-430)
# ---------------------------------------------------------------------------
def test_topology2d_integrate_matches_fft_closed_form(orc):
    rng = np.random.default_rng(81)
    shape = (7, 6)
    vs = np.array([1.2, 0.9])
    g = rng.standard_normal(shape + (2, 2))
    m = np.array([2.0, -1.5])
    got = orc.topo2_integrate(g, m, shape, vs)
    E = [np.exp(2j * np.pi * np.fft.fftfreq(shape[0]))[:, None],
         np.exp(2j * np.pi * np.fft.fftfreq(shape[1]))[None, :]]
    den = sum(2 - 2 * np.real(e) for e in E)
    den[0, 0] = 1.0
    for n in range(2):
        U = sum(np.conj(E[a] - 1) * np.fft.fft2(g[..., n, a]) for a in range(2)) / den
        U[0, 0] = 0.0
        exp = np.real(np.fft.ifft2(U)) * vs[n] + m[n]
        assert np.allclose(got[..., n], exp, atol=1e-12)


def test_topology2d_gradients_then_integrate_roundtrip(orc):
    rng = np.random.default_rng(82)
    shape = (9, 8)
    vs = np.array([0.8, 1.3])
    w = rng.standard_normal(shape + (2,)) * 3
    g = orc.topo2_gradients(w, shape, vs)
    back = orc.topo2_integrate(g, w.reshape(-1, 2).mean(0), shape, vs)
    assert np.allclose(back, w, atol=1e-10)


@pytest.mark.parametrize("fold", [False, True])
def test_topology2d_limit_single_corner(orc, fold):
    """As in 3D: the limited corner's two columns are the edge differences
    that start at the cell's lower face, their determinant is the corner
    Jacobian, and they become (1-a)J + aI one step past the first blend in
    range."""
    rng = np.random.default_rng(83)
    shape = (6, 7)
    vs = np.ones(2)
    e1, e2 = 0.01, 100.0
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    w = pos + rng.normal(0, 0.05, pos.shape)
    cell = (2, 3)
    if fold:
        w[3, 4, 0] -= 1.6
    Jc, _ = orc.corner_jacobians_2D(w.ravel(), *shape, vs, e1, e2)
    Jc = Jc.reshape(shape[0] - 1, shape[1] - 1, 4)
    g0 = orc.topo2_gradients(w, shape, vs)
    vertices = [(0, 0), (1, 0), (0, 1), (1, 1)]
    corners = [int(np.argmin(Jc[cell]))] if fold else range(4)
    for c in corners:
        jc = np.full_like(Jc, 1.0)
        jc[cell + (c,)] = Jc[cell + (c,)] if fold else -1.0
        g = orc.topo2_limit(g0, jc.ravel(), shape, e1, e2)
        locs = []
        for a in range(2):
            v = [cell[k] + vertices[c][k] for k in range(2)]
            v[a] = cell[a]
            locs.append(tuple(v))
        M = np.stack([g0[locs[a]][:, a] for a in range(2)], 1)
        assert np.isclose(np.linalg.det(M), Jc[cell + (c,)], rtol=1e-12, atol=1e-14)
        alpha, det = 0.0, np.linalg.det(M)
        while det < e1 or det > e2:
            alpha = min(alpha + 0.1, 1.0)
            det = np.linalg.det((1 - alpha) * M + alpha * np.eye(2))
        alpha = min(alpha + 0.1, 1.0)
        new = (1 - alpha) * M + alpha * np.eye(2)
        expect = g0.copy()
        for a in range(2):
            expect[locs[a]][:, a] = new[:, a]
        assert np.allclose(g, expect, atol=1e-14, rtol=0)


def test_topology2d_invariants(orc):
    from binder_amd import synth
    shape = (24, 20)
    rng = np.random.default_rng(84)
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    f = pos + synth.dct_displacement(shape, 2.5, rng, ncoef=4)
    t = orc.topology_correction_2D()
    Jc0, cnt0 = orc.corner_jacobians_2D(f.ravel(), *shape, np.ones(2), 0.01, 100)
    assert cnt0[0] > 0  # the field folds
    out = t.correct_topology(f.ravel(), 0.01, 100, 10, 10, 1, np.ones(2), *shape)
    Jc, cnt = orc.corner_jacobians_2D(out.ravel(), *shape, np.ones(2), 0.01, 100)
    assert cnt[0] == 0 and cnt[1] == 0 and Jc.min() >= 0.01
    assert np.allclose(out.reshape(-1, 2).mean(0), f.reshape(-1, 2).mean(0), atol=1e-10)
    same = t.correct_topology(pos.ravel(), 0.01, 100, 10, 10, 1, np.ones(2), *shape)
    assert np.array_equal(same, pos) and t.last_stats[0] == 0
