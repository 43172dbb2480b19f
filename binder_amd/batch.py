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
