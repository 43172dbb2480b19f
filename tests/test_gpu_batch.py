"""Lockstep batched EM over several registrations (binder_batch_em) against
each registration's own EM (BINDER/src/cpu/Registration.cpp:122-217): the
same fields, theta, objective traces and iteration counts, bit for bit, for
uneven shapes, different iteration caps and registrations that stop at
different iterations; and the level-synchronous batched pyramid against the
plain pyramid of each pair (BINDER/FastRegistration.py:443-546)."""
import numpy as np
import pytest

from binder_amd import synth

pytestmark = pytest.mark.gpu


def _problems():
    out = []
    for shape, seed in (((24, 20, 18), 1), ((30, 16, 22), 2), ((17, 26, 20), 3)):
        pair = synth.make_pair(shape, 4, 1.5, seed=seed, semi=(3.0, 8.0))
        out.append(synth.problem_from_pair(pair, gamma=1e-4))
    return out


@pytest.mark.parametrize("update,caps,th", [(1, [40, 7, 25], 1e-5), (1, [300, 300, 300], 3e-5),
                                            (0, [40, 7, 25], 1e-5)])
def test_batch_em_equals_single_em(core, update, caps, th):
    print(f"batched EM vs single EM, update={update}, caps {caps}, th {th}", flush=True)
    probs = _problems()
    alone = []
    for prob, cap in zip(probs, caps):
        r = synth.build_registration(core, prob)
        r.EM(cap, th, update, 0)
        alone.append((r.get_final_locations_as_numpy_array(), r.get_objective_trace(),
                      np.asarray(r.get_likelihood_parameters()),
                      np.asarray(r.get_dream_locations())))
    regs = [synth.build_registration(core, prob) for prob in probs]
    its = core.Batch(regs).EM(caps, th, update)
    print(f"iterations alone {[len(a[1]) for a in alone]}, batched {list(its)}", flush=True)
    for r, a, n in zip(regs, alone, its):
        assert n == len(a[1])
        assert np.array_equal(r.get_final_locations_as_numpy_array(), a[0])
        assert np.array_equal(r.get_objective_trace(), a[1])
        assert np.array_equal(np.asarray(r.get_likelihood_parameters()), a[2])
        assert np.array_equal(np.asarray(r.get_dream_locations()), a[3])
    # pairs leave the lockstep loop at different iterations (caps or the
    # convergence rule)
    if caps[0] == caps[1]:
        assert min(its) < caps[0]


def test_batch_em_zero_cap_and_rerun(core):
    """A zero cap leaves its registration untouched; a second batched call
    continues from where the first left every registration, as a second
    single EM call does."""
    print("batched EM: zero cap and a second call", flush=True)
    probs = _problems()
    ref = []
    for prob, caps in zip(probs, ((5, 4), (0, 6), (3, 3))):
        r = synth.build_registration(core, prob)
        for c in caps:
            r.EM(c, -1e30, 1, 0)
        ref.append(r.get_final_locations_as_numpy_array())
    regs = [synth.build_registration(core, prob) for prob in probs]
    b = core.Batch(regs)
    assert list(b.EM([5, 0, 3], -1e30, 1)) == [5, 0, 3]
    assert list(b.EM([4, 6, 3], -1e30, 1)) == [4, 6, 3]
    for r, f in zip(regs, ref):
        assert np.array_equal(r.get_final_locations_as_numpy_array(), f)


def test_batch_rejects_mixed_bins(core):
    probs = _problems()
    a = synth.build_registration(core, probs[0])
    b = synth.build_registration(core, probs[1])
    K = 8
    b.setMILikelihood(2.0, np.full(K * K, 1.0 / K), K, K, 1,
                      np.minimum(probs[1].binned_nodes, K - 1).ravel(),
                      np.minimum(probs[1].binned_voxels, K - 1).ravel())
    with pytest.raises(ValueError):
        core.Batch([a, b])
    with pytest.raises(ValueError):
        core.Batch([a, a])


def test_batched_pyramid_equals_single_pyramids(core):
    """Three uneven pairs through the level-synchronous batched pyramid
    (topology correction included): per-level iteration counts and final
    fields identical to each pair's own pyramid."""
    from binder_amd import batch
    print("batched pyramid vs single pyramids", flush=True)
    specs = [{"shape": s, "rms": r, "seed": 20 + i}
             for i, (s, r) in enumerate((((48, 40, 44), 2.0), ((40, 56, 36), 3.0),
                                         ((64, 48, 40), 2.5)))]
    pairs = [synth.make_pair(sp["shape"], 5, sp["rms"], sp["seed"], semi=(3.0, 12.0))
             for sp in specs]
    frs, _ = batch.register_pairs_batched(pairs, workers=3)
    for fr, pair in zip(frs, pairs):
        one = batch.register_one(pair)
        assert [lv[1] for lv in fr.level_log] == [lv[1] for lv in one.level_log]
        assert np.array_equal(fr.deformedVoxelPositions, one.deformedVoxelPositions)
    print(f"iterations {[[lv[1] for lv in fr.level_log] for fr in frs]}", flush=True)
