"""The launch-shape autotuner (no reference counterpart: the reference fixes
its launch shapes, BINDER/src/gpu/Registration.cu:756-761,
BINDER/src/gpu/NonLinearTransformation.cu:341-345): EM and Gibbs sweeps at
96x115x96 (large enough for every tunable: E-step grids above 2^18 voxels,
persistent y/z passes whose matrices leave room for two blocks per CU) with
default launch shapes, with every tunable forced to a
non-default shape, and with timed tuning give bit for bit the same fields,
objective traces, histograms and sampler moments."""
import numpy as np
import pytest

from binder_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob():
    return synth.config_problem("c2", seed=0, scale=0.6)[1]


def _run(core, prob):
    r = synth.build_registration(core, prob)
    r.EM(4, -1e30, 1, 0)
    out = [r.get_final_locations_as_numpy_array(), r.get_objective_trace()]
    mll, H = r.e_step(return_histogram=True)
    out += [np.array([mll]), H]
    mean, cov = r.sampler(3, 1, 3, 0, 1, 1, 1, 1)
    out += [mean, cov]
    return out


def test_autotuned_shapes_give_default_results(core, prob):
    print("autotuner: default vs forced non-default vs tuned launch shapes", flush=True)
    core.autotune(0)
    core.autotune_clear()
    ref = _run(core, prob)
    for mode in (2, 1):
        core.autotune_clear()
        core.autotune(mode)
        try:
            got = _run(core, prob)
            rep = core.autotune_report()
        finally:
            core.autotune(0)
        ops = {row[0] for row in rep}
        print(f"mode {mode}: {rep}", flush=True)
        assert {"estep_em", "estep_sample", "contract"} <= ops, ops
        if mode == 2:
            assert all(row[2] != row[3] for row in rep)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
    core.autotune_clear()
