"""Independent Gibbs chains with graph-captured sweeps
(Registration.sampler_chains): each chain draws exactly what
sampler_run with its seed draws, the pooled accumulators equal the pairwise
(Chan et al.) merge of the per-chain results up to rounding, graph replays
equal eager launches bit for bit, and the handle's state is restored
afterwards."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem():
    from binder_amd import synth
    pair = synth.make_pair((24, 20, 18), 4, 1.5, seed=5, semi=(3.0, 8.0))
    return synth.problem_from_pair(pair, gamma=1e-4)


def _handle(core, prob):
    from binder_amd import synth
    r = synth.build_registration(core, prob)
    r.EM(3, -1e30, 1, 0)
    return r


def _merge(parts):
    """Chan et al. pooled mean / M2 of [(n, mean, M2)] in order."""
    n, mean, m2 = parts[0]
    mean, m2 = mean.copy(), m2.copy()
    for nb, mb, m2b in parts[1:]:
        tot = n + nb
        d = mb - mean
        mean = mean + d * (nb / tot)
        m2 = m2 + m2b + np.einsum("...i,...j->...ij", d, d) * (n * nb / tot)
        n = tot
    return n, mean, m2


@pytest.mark.parametrize("Nb,Ns", [(2, 5), (3, 14), (3, 31)])
def test_chains_match_per_seed_runs_and_graph_replays(core, Nb, Ns):
    # (3, 14) and (3, 31): 17 and 34 sweeps, the 16-slot field ring wraps
    # within a chain, once and twice
    print(f"sampler chains 24x20x18: 3 chains x ({Nb} burn-in + {Ns}) sweeps", flush=True)
    prob = _problem()
    seeds = [11, 12, 13]
    # one chain per seed with the plain sampler, each from the same state
    parts, covs = [], []
    for sd in seeds:
        r = _handle(core, prob)
        r.sampler_run(sd, Nb, Ns)
        mean, cov = r.sampler_result()
        parts.append((Ns, mean, cov * (Ns - 1)))
        covs.append(cov)
    n, mref, m2ref = _merge(parts)
    covref = m2ref / (n - 1)
    r = _handle(core, prob)
    start = r.get_final_locations_as_numpy_array()
    theta = np.asarray(r.get_likelihood_parameters())
    r.sampler_chains(seeds, Nb, Ns, use_graph=False)
    me, ce = r.sampler_result()
    np.testing.assert_allclose(me, mref, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(ce, covref, rtol=1e-9, atol=1e-9)
    # state restored
    np.testing.assert_array_equal(r.get_final_locations_as_numpy_array(), start)
    np.testing.assert_array_equal(np.asarray(r.get_likelihood_parameters()), theta)
    # graph replays draw exactly what eager launches draw
    r.sampler_chains(seeds, Nb, Ns, use_graph=True)
    mg, cg = r.sampler_result()
    np.testing.assert_array_equal(mg, me)
    np.testing.assert_array_equal(cg, ce)
    # a single chain is the plain sampler
    r.sampler_chains([12], Nb, Ns, use_graph=True)
    m1, c1 = r.sampler_result()
    np.testing.assert_array_equal(m1, parts[1][1])
    np.testing.assert_array_equal(c1, covs[1])
    # the graph is replayed several times with eager work in between
    for _ in range(2):
        r.EM(1, -1e30, 1, 0)
        r.sampler_chains(seeds, Nb, Ns, use_graph=True)
        mg, cg = r.sampler_result()
        r.sampler_chains(seeds, Nb, Ns, use_graph=False)
        me, ce = r.sampler_result()
        np.testing.assert_array_equal(mg, me)
        np.testing.assert_array_equal(cg, ce)
        assert np.all(np.isfinite(ce))
