"""Many registration pairs and many Gibbs chains on one GPU.

Pairs of uneven shapes register concurrently: each pair's pyramid
(FastRegistration.run) runs in its own host thread with its own
registration handles, so its HIP streams carry their kernels side by side
with the other pairs' (ctypes releases the GIL during every backend call).
Then each pair runs independent Gibbs chains from its registered state
(Registration.sampler_chains: one captured HIP graph of a chain's sweeps,
replayed per chain with a different Philox key), again several pairs at a
time.  The reference registers one pair per process and runs one chain per
sampler() call (BINDER/FastRegistration.py:443-750).
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import synth

AXES = (96, 112, 128, 144, 160, 192)


def batch_specs(n=48, seed=0):
    """n pairs: each axis uniform on AXES, warp RMS uniform on [2, 6] voxels,
    the 3D workload generator otherwise (12 ellipsoids, 4 tissue classes,
    x^0.7 intensity map, 32 bins)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shape = tuple(int(a) for a in rng.choice(AXES, 3))
        out.append({"shape": shape, "rms": float(rng.uniform(2.0, 6.0)), "seed": 10_000 + i})
    return out


def make_pair(spec):
    return synth.make_pair(spec["shape"], 12, spec["rms"], spec["seed"])


def register_one(pair, **kw):
    from .FastRegistration import FastRegistration
    fr = FastRegistration(voxels=pair.fixed, nodes=pair.moving,
                          **dict(synth.REGISTRATION_KW, **kw))
    fr.run(list(synth.PYRAMID), synth.EM_ITERATIONS)
    return fr


def register_pairs(pairs, workers=8, **kw):
    """Registers every pair (plain pyramid), `workers` pairs at a time;
    returns the FastRegistration objects in order and the wall-clock."""
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        regs = list(ex.map(lambda p: register_one(p, **kw), pairs))
    return regs, time.perf_counter() - t0


def register_pairs_batched(pairs, workers=8, **kw):
    """Registers every pair with the plain pyramid, all pairs' EM iterations
    sharing launches: one Batch slot per pair runs the pair's current level;
    when a pair's level converges, a worker thread finishes that level
    (topology correction on the GPU) and sets up the next one while the other
    slots keep iterating, and the next level rejoins the batch.  Each
    iteration runs the E-steps of all running pairs in one launch over
    (pair, block) items, their theta steps in one launch, and the smoothing
    passes of small grids as one launch per pass.  Every pair ends with
    exactly the field and iteration counts of register_one.  Returns the
    FastRegistration objects in order and the wall-clock."""
    from concurrent.futures import FIRST_COMPLETED, wait
    from .FastRegistration import FastRegistration, Registration
    t0 = time.perf_counter()
    n = len(pairs)
    frs, gens = [None] * n, [None] * n
    kw = dict(synth.REGISTRATION_KW, **kw)
    th = kw.get("convergence_th", 5e-5)
    upd = int(kw.get("update_likelihood_parameters", 1))
    bt = Registration.Batch(n)

    stats = {"runs": 0, "idle_waits": 0, "run_s": 0.0, "wait_s": 0.0, "setup_s": 0.0,
             "slot_iterations": 0}

    def timed(fn):
        t = time.perf_counter()
        out = fn()
        stats["setup_s"] += time.perf_counter() - t
        return out

    def first(i):
        frs[i] = FastRegistration(voxels=pairs[i].fixed, nodes=pairs[i].moving, **kw)
        gens[i] = frs[i].pyramid_levels(list(synth.PYRAMID), synth.EM_ITERATIONS)
        return next(gens[i], None)

    with ThreadPoolExecutor(max_workers=workers) as ex:
        pending = {ex.submit(timed, lambda i=i: first(i)): i for i in range(n)}
        running = 0
        while pending or running:
            if running:
                ready = [f for f in pending if f.done()]
            else:
                stats["idle_waits"] += 1
                tw = time.perf_counter()
                ready = list(wait(list(pending), return_when=FIRST_COMPLETED).done)
                stats["wait_s"] += time.perf_counter() - tw
            for f in ready:
                i = pending.pop(f)
                step = f.result()
                if step is not None:
                    bt.set(i, step[0], step[1], upd)
                    running += 1
            if running:
                tr = time.perf_counter()
                fin = bt.run(th, 1)
                stats["run_s"] += time.perf_counter() - tr
                stats["runs"] += 1
                for i in np.nonzero(fin >= 0)[0]:
                    running -= 1
                    bt.set(int(i), None, 0, upd)
                    pending[ex.submit(timed, lambda i=int(i): next(gens[i], None))] = int(i)
    stats = {k: round(v, 3) if isinstance(v, float) else v for k, v in stats.items()}
    register_pairs_batched.stats = stats
    return frs, time.perf_counter() - t0


def full_resolution_handle(fr):
    """The registration handle holding the pair's registered state: the
    pyramid's last (full-resolution) level, whose device field is the
    topology-corrected result and whose theta is the fitted one -- the state
    FastRegistration.sample_param starts from, without re-uploading it."""
    if fr.current_resolution != 1:
        raise ValueError("the pyramid did not end at full resolution")
    return fr.registratorObj


def sweeps_for(site_updates, nv):
    """Sweeps of one chain for a given number of site updates (one site
    update = one voxel's latent resample)."""
    return max(1, int(math.ceil(float(site_updates) / nv)))


def run_chains(fr, chains=16, site_updates=1e6, seed0=1, use_graph=True, keep=False):
    """`chains` independent chains of `site_updates` site updates each from
    the pair's registered state; returns (voxel updates, seconds[, handle])."""
    h = full_resolution_handle(fr)
    nv = int(np.prod(fr.voxels_shape))
    ns = sweeps_for(site_updates, nv)
    seeds = [seed0 + c for c in range(chains)]
    t0 = time.perf_counter()
    h.sampler_chains(seeds, 0, ns, use_graph=use_graph)
    dt = time.perf_counter() - t0
    out = (chains * ns * nv, dt)
    return out + (h,) if keep else out


def chains_for_pairs(regs, chains=16, site_updates=1e6, workers=8, use_graph=True):
    """Chains for every pair, `workers` pairs at a time: (total voxel
    updates, wall-clock)."""
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=workers) as ex:
        res = list(ex.map(lambda r: run_chains(r, chains, site_updates, use_graph=use_graph)[0],
                          regs))
    return int(sum(res)), time.perf_counter() - t0
