"""Linear (order 1) and quadratic (order 2) B-spline node weights on the HIP
E-steps against the FP64 oracle, whose spline weights are pinned against the
reference's golden vectors (test_oracle.py).

Reference: BINDER/src/cpu/Spline.cpp:8-110 (weights and node offsets per
order), Registration.cpp:219-330 (E-step), :332-440 (log-space variant),
FastRegistration.py:409-427 (spline_order), Registration.cpp:31-45 (the
order's smoothing sigma, Smoothing.cpp).
"""
import numpy as np
import pytest
from scipy import stats

from binder_amd import synth
from binder_amd.utils import initialize_fourier_domain_smoothing

from test_gpu_estep_em import _random_problem


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("shape,K,L,oob", [
    ((12, 10, 9), 8, 6, False),
    ((17, 13, 11), 32, 32, True),
    ((15, 12, 10), 48, 40, False),   # K > 32: the FP64-partials E-step
    ((31, 23, 1), 16, 8, False),
    ((40, 33, 1), 32, 32, True),
])
def test_estep_orders_match_oracle(core, orc, order, shape, K, L, oob):
    print(f"estep order={order} shape={shape} K={K} L={L} oob={oob}", flush=True)
    prob = _random_problem(shape, K, L, seed=sum(shape) + K + order, oob=oob)
    g = synth.build_registration(core, prob, spline_order=order)
    o = synth.build_registration(orc, prob, spline_order=order)
    mll_g, H_g = g.e_step(return_histogram=True)
    mll_o, H_o = o.e_step(return_histogram=True)
    assert abs(mll_g - mll_o) <= 1e-6 * abs(mll_o)
    assert abs(H_g.sum() - np.prod(shape)) <= 1e-6 * np.prod(shape)
    assert np.max(np.abs(H_g - H_o)) <= 1e-5 * max(1.0, H_o.max())
    dg = np.asarray(g.get_dream_locations()).reshape(prob.final_locations.shape)
    do = np.asarray(o.get_dream_locations()).reshape(prob.final_locations.shape)
    assert np.max(np.abs(dg - do)) <= 2e-4


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2])
def test_estep_orders_on_nodes(core, order):
    """Positions exactly on integer (order 1) or half-integer (order 2)
    coordinates: the node weights are then {1, 0} or {1/2, 1/2}, so the
    dream location of an in-grid voxel under a uniform theta is the position
    itself."""
    print(f"estep order={order} on grid points", flush=True)
    shape = (10, 9, 8)
    prob = _random_problem(shape, 6, 5, seed=13)
    prob.theta = np.full((6, 5), 0.2)
    prob.final_locations = prob.voxel_pos + (0.5 if order == 2 else 0.0)
    prob.final_locations[..., 0] += 0.25 * (order == 1)
    g = synth.build_registration(core, prob, spline_order=order)
    mll, H = g.e_step(return_histogram=True)
    assert abs(mll - np.prod(shape) * np.log(5.0)) <= 1e-6 * abs(mll)
    dream = np.asarray(g.get_dream_locations()).reshape(prob.final_locations.shape)
    assert np.allclose(dream, prob.final_locations, atol=2e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("shape,iters", [((20, 18, 16), 4), ((48, 40, 1), 5)])
def test_em_orders_match_oracle(core, orc, order, shape, iters):
    print(f"EM order={order} shape={shape} iters={iters}", flush=True)
    pair = synth.make_pair(shape[:2] if shape[2] == 1 else shape, 4, 1.5, seed=7,
                           semi=(3.0, 8.0))
    prob = synth.problem_from_pair(pair, gamma=1e-4)
    g = synth.build_registration(core, prob, spline_order=order)
    o = synth.build_registration(orc, prob, spline_order=order)
    g.EM(iters, -1e30, 1, 0)
    o.EM(iters, -1e30, 1, 0)
    tg, to = g.get_objective_trace(), o.get_objective_trace()
    assert tg.shape == to.shape == (iters, 4)
    assert np.all(np.abs(tg - to) <= 1e-5 * np.abs(to) + 1e-6)
    fg = g.get_final_locations_as_numpy_array().astype(np.float64)
    fo = o.get_final_locations_as_numpy_array()
    assert np.max(np.abs(fg - fo)) <= 1e-3
    thg = np.asarray(g.get_likelihood_parameters())
    tho = np.asarray(o.get_likelihood_parameters())
    assert np.max(np.abs(thg - tho)) <= 1e-5


def _weights(order, t):
    """B-spline weights of `order` at fractional offset t and the offsets of
    the nodes they belong to, from the spline's definition."""
    if order == 1:
        return np.array([1 - t, t]), np.array([0, 1])
    return np.array([0.5 * (0.5 - t) ** 2, 0.75 - t ** 2, 0.5 * (0.5 + t) ** 2]), \
        np.array([-1, 0, 1])


def _posterior(prob, idx, order):
    p = prob.final_locations[idx]
    base = np.floor(p + (0.5 if order == 2 else 0.0)).astype(int)
    wo = [_weights(order, p[a] - base[a]) for a in range(3)]
    nodes, probs = [], []
    for c in np.ndindex(*(len(w) for w, _ in wo)):
        node = base + np.array([wo[a][1][c[a]] for a in range(3)])
        inside = all(0 <= node[a] < prob.shape[a] for a in range(3))
        k = prob.binned_nodes[tuple(node)] if inside else 0
        probs.append(np.prod([wo[a][0][c[a]] for a in range(3)]) *
                     prob.theta[k, prob.binned_voxels[idx]])
        nodes.append(tuple(int(v) for v in node))
    probs = np.array(probs)
    return nodes, probs / probs.sum()


def _check_draws(impl, order):
    rng = np.random.default_rng(4)
    shape, K, L = (8, 7, 6), 6, 5
    pos = np.stack(np.indices(shape, dtype=np.float64), -1)
    theta = rng.random((K, L)) + 0.1
    theta /= theta.sum(1, keepdims=True)
    D = initialize_fourier_domain_smoothing(list(shape), True, 2, 3).reshape(shape)
    g = 0.05 * np.prod(shape)
    prob = synth.Problem(shape=shape, final_locations=pos + rng.normal(0, 0.8, pos.shape),
                         voxel_pos=pos, binned_nodes=rng.integers(0, K, shape).astype(np.int32),
                         binned_voxels=rng.integers(0, L, shape).astype(np.int32),
                         theta=theta, D=D, gammas=(g, g, g), K=K, L=L, alpha=2.0,
                         nnz=int(np.prod(shape)))
    r = synth.build_registration(impl, prob, spline_order=order)
    voxels = [(1, 2, 3), (4, 4, 2), (6, 1, 5), (0, 0, 0)]
    S = 3000
    counts = [dict() for _ in voxels]
    for s in range(S):
        r.sampler(500 + s, 0, 2, 0, 1, 1, 0, 0)
        dream = np.asarray(r.get_dream_locations()).reshape(prob.final_locations.shape)
        for c, v in zip(counts, voxels):
            key = tuple(int(round(x)) for x in dream[v])
            c[key] = c.get(key, 0) + 1
    for c, v in zip(counts, voxels):
        nodes, probs = _posterior(prob, v, order)
        obs = np.array([c.get(n, 0) for n in nodes], float)
        assert obs.sum() == S, (v, c)  # every draw is a neighbour of positive weight
        idx = np.argsort(probs)[::-1]
        o, e = obs[idx], probs[idx] * S
        k = max(1, min(len(e) - 1, np.count_nonzero(e >= 5)))
        o = np.append(o[:k], o[k:].sum())
        e = np.append(e[:k], e[k:].sum())
        chi2 = float(np.sum((o - e) ** 2 / np.maximum(e, 1e-300)))
        assert stats.chi2.sf(chi2, len(o) - 1) > 1e-4, (v, chi2, len(o))


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2])
def test_estep_draw_frequencies_orders(core, order):
    """HIP categorical node draws with order-1 / order-2 weights follow the
    exact posterior over the 8 / 27 neighbours (chi-square)."""
    print(f"E-step draw frequencies order={order}", flush=True)
    _check_draws(core, order)


@pytest.mark.parametrize("order", [1, 2])
def test_estep_draw_frequencies_orders_oracle(orc, order):
    """The oracle's node draws with order-1 / order-2 weights follow the same
    exact posterior (pins the posterior helper and the oracle's offsets)."""
    _check_draws(orc, order)


@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2])
def test_fastregistration_order_vs_oracle(core, orc, monkeypatch, order):
    """The pyramid with spline_order 1 / 2 (the order's smoothing sigma and
    node weights at every level), HIP backend against the oracle backend."""
    import binder_amd.FastRegistration as fr
    print(f"FastRegistration spline_order={order} pyramid [2, 1] x 3 iterations", flush=True)
    pair = synth.make_pair((40, 36, 32), 6, 2.0, 4, semi=(4.0, 12.0), bins=32)

    def run(backend):
        monkeypatch.setattr(fr, "Registration", backend)
        reg = fr.FastRegistration(voxels=pair.fixed, nodes=pair.moving, threads=8,
                                  num_voxel_intesity_levels=32, num_node_intensity_levels=32,
                                  gamma=1e-5, topology_correction=False, spline_order=order)
        reg.run([2, 1], 1000, level_iterations=[3, 3])
        return np.asarray(reg.deformedVoxelPositions, np.float64)

    fg, fo = run(core), run(orc)
    print(f"  field max {np.abs(fg - fo).max():.2e}", flush=True)
    assert np.abs(fg - fo).max() <= 2e-3
