"""Benchmark of the registration-and-sampling path on one MI355X.

Workload: the 3D MI registration of 160x192x160 (synthetic 12-ellipsoid
phantom, 6-voxel-RMS DCT warp, x^0.7 intensity map, 32 bins): EM
registration from a cold start to convergence, topology correction of the
result, then Gibbs sampler sweeps.  A step is one Gibbs sweep over all
Nv = 4 915 200 voxels (one site update per voxel: categorical E-draw,
Dirichlet theta draw, DCT-domain deformation draw, FP64 mean/covariance
update).  value = Nv * steps / time, inputs resident in HBM.

Prints one JSON line.  `--gpus N` with N > 1 is refused (single-GPU build).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("3D reg time-to-tolerance (warm-start vs plain) + Gibbs voxel-updates/s, 1 GPU")
HBM_PEAK_GBS = 8000.0
WELFORD_BATCH = 16  # samples folded per accumulator pass (kWelfordBatch)
# dense BF16 MFMA peak of MI355X: the DCT passes run split-BF16 MFMA
# (three products per algorithmic multiply-add, FP32 accumulation)
BF16_PEAK_TFLOPS = 2500.0
# minimum HBM bytes of one Gibbs site update at the reference API's dtypes
# (f32 field, int32 bins): E-draw 32 B + deformation draw 40 B + FP64
# Welford read/write of the mean (2 x 24 B) and the full 3x3 M2 (2 x 72 B)
SWEEP_BYTES_PER_UPDATE = 264

# HBM bytes per launch at the 160x192x160 workload, from two rocprofv3 PMC
# passes (FETCH_SIZE, WRITE_SIZE; KiB per dispatch) scaled by the readings of
# a plain 1 GiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  Commands and
# per-kernel table: profiles/gibbs_sweep_160x192x160_hbm_traffic.md.
MEASURED_TRAFFIC = {
    (160, 192, 160): {
        # 2048 blocks, the default shape the bench's tuner keeps
        "estep_sample": (46740.0 * 2.0 + 74001.5) * 1024,
    },
}


def kernel_model(name, nv, nn, N, dims):
    """Algorithmic HBM bytes and flops of one launch of `name` at this size."""
    nx, ny, nz = dims
    f32 = 4
    if name in ("estep", "estep_sample"):
        # read d (N f32) + fixed bin (int32) + moving bins (int32 per node,
        # each node read once), write dream (N f32): 32 B per voxel in 3D.
        # The kernels read u8 bins (26 B per voxel): packed_bin_bytes().
        return nv * (N * f32 + 4 + N * f32) + nn * 4, 0.0
    if name == "dct_x_fused":
        # read the coefficients and D (once per component), write the field;
        # forward and inverse transform along x
        return nv * N * 3 * f32, 2 * 2.0 * nx * nv * N
    if name.startswith("dct_"):
        # five-pass schedule (3D): z forward reads dream and x, the last
        # pass (y inverse) reads x for the final addition
        axis = {"x": nx, "y": ny, "z": nz}[name[4]]
        extra = nv * N * f32 if name in ("dct_z_fwd", "dct_y_inv") else 0
        return nv * N * 2 * f32 + extra, 2.0 * axis * nv * N
    if name == "welford":
        # one launch folds WELFORD_BATCH sweeps' samples into the FP64
        # accumulators: read the samples, read and write mean and M2 once
        npair = N * (N + 1) // 2
        return nv * (WELFORD_BATCH * N * f32 + 2 * 8 * N + 2 * 8 * npair), 0.0
    return None, None


def packed_bin_bytes(nv, nn, N):
    """The E-steps' own minimum: u8 fixed and moving bins instead of int32."""
    return nv * (2 * N * 4 + 1) + nn


def host_threads():
    """Host cores this process may use (affinity, capped by OMP_NUM_THREADS)."""
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    return min(n, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else n


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


from binder_amd.synth import REGISTRATION_KW as REG_KW  # noqa: E402
TOLERANCE_VOX_RMS = 0.1


def _rms(a, b):
    return float(np.sqrt(np.mean(np.sum((np.asarray(a, np.float64) - b) ** 2, -1))))


# runs of each timed (stopped) pipeline; the median is reported
TIMED_REPEATS = 5


def registration(args, pair, net=None, ref=None, ref_objective=None):
    """Pyramid registration as run_BINDER runs it -- from the identity (the
    plain cold start) or, with `net`, from the learned warm start
    (FastRegistration.run_warm) -- and its time to tolerance.

    The pipeline runs to convergence (A), then its last level is retraced in
    chunks of 5 EM iterations (B; deterministic, so B follows A; each
    chunk's field is topology-corrected as the pipeline would return it).
    Each chunk is judged by up to three criteria:
      * oracle field: within 0.1 voxel RMS of `ref` (the FP64 oracle's
        converged field of the same pipeline: the time-to-tolerance target);
      * objective: mlp at or below `ref_objective` (the cold start's final
        objective, used to judge the warm start);
      * own field: within 0.1 voxel RMS of A's own converged field.
    For every criterion that is reached, the pipeline is timed end to end
    (C) stopped at that iteration: the median of TIMED_REPEATS runs."""
    from binder_amd.FastRegistration import FastRegistration

    warm = net is not None
    if warm:
        levels = list(args.warm_levels)
    else:
        levels = [2, 1] if args.quick else list(args.resolutions)
    iters = 3 if args.quick else args.em_iters
    kw = dict(REG_KW, convergence_th=args.convergence_th)
    if args.quick:
        kw["max_outer_it"] = 0

    def new():
        return FastRegistration(voxels=pair.fixed, nodes=pair.moving, **kw)

    def run(fr, level_iterations=None):
        if warm:
            return fr.run_warm(net, levels, iters, level_iterations=level_iterations)
        fr.run(levels, iters, level_iterations=level_iterations)
        return True

    # A: to convergence (the reference's "Time for EM")
    A = new()
    t0 = time.perf_counter()
    used = run(A)
    conv_s = time.perf_counter() - t0
    F = np.asarray(A.deformedVoxelPositions, np.float64)
    ran = [int(lv[0]) for lv in A.level_log]
    counts = [int(lv[1]) for lv in A.level_log]
    final_obj = float(A.registratorObj.get_objective_trace()[-1, 3])
    log(f"{'warm' if warm else 'cold'} pyramid {ran}: iterations {counts} in {conv_s:.3f}s")
    out = {"resolutions": ran, "converged_s": round(conv_s, 4),
           "converged_iterations": counts,
           "level_seconds": [round(lv[2], 4) for lv in A.level_log],
           "final_objective": round(final_obj, 2)}
    if warm:
        out["warm_start_used"] = bool(used)
        out["start_objectives_warm_identity"] = [round(v, 2) for v in A.warm_objectives]
    if ref is not None:
        out["converged_rms_vs_oracle_vox"] = round(_rms(F, ref), 4)
    if ref_objective is not None:
        out["final_objective_minus_reference"] = round(final_obj - ref_objective, 2)
    if args.quick:
        return A, F, out, F
    # B: retrace the last level in chunks (untimed)
    B = new()
    if warm and used:
        # the first level starts from the theta the fallback check fitted
        B.theta = A.warm_theta.copy()
    if warm and used and len(ran) == 1:
        B.move_resolution(ran[0], resize_field=False)
        reg = B._warm_registration(A.warm_coef, ran[0])
    else:
        if warm and used:
            B.run_warm(net, ran[:-1], iters, level_iterations=counts[:-1], compare_iterations=0)
        else:
            B.run(ran[:-1], iters, level_iterations=counts[:-1])
        B.move_resolution(ran[-1])
        reg = B._new_registration(np.ascontiguousarray(B.deformedVoxelPositions))
    reg.setMILikelihood(B.alpha, np.ascontiguousarray(B.theta).ravel(),
                        B.num_node_intensity_levels, B.num_voxel_intensity_levels, 1,
                        np.ascontiguousarray(B.binned_nodes).ravel(),
                        np.ascontiguousarray(B.binned_voxels).ravel())
    crit = {"oracle_field": None, "objective": None, "own_field": None}
    trace = []
    k, chunk = 0, 5
    while k < counts[-1]:
        step = min(chunk, counts[-1] - k)
        reg.EM(step, -1e30, 1, 0)
        k += step
        obj = float(reg.get_objective_trace()[-1, 3])
        f = reg.get_final_locations_as_numpy_array()
        if B.top_cor_obj:  # compare what the pipeline would return here
            f = B.correct_topology(f)
        row = {"k": k, "objective": round(obj, 2), "rms_own": round(_rms(f, F), 4)}
        if ref is not None:
            row["rms_oracle"] = round(_rms(f, ref), 4)
            if crit["oracle_field"] is None and row["rms_oracle"] < TOLERANCE_VOX_RMS:
                crit["oracle_field"] = k
        if ref_objective is not None and crit["objective"] is None and obj <= ref_objective:
            crit["objective"] = k
        if crit["own_field"] is None and row["rms_own"] < TOLERANCE_VOX_RMS:
            crit["own_field"] = k
        trace.append(row)
    out["retrace"] = trace
    # C: the timed pipeline, stopped at each criterion's iteration, and once
    # to convergence (A's time includes first-use costs: buffers, tuning)
    tol = {}
    timed = {}

    def timed_run(level_iterations):
        # median wall-clock of TIMED_REPEATS runs of the stopped pipeline
        # (each from a fresh FastRegistration, made outside the timed region)
        secs, Fc = [], None
        for _ in range(TIMED_REPEATS):
            C = new()
            t0 = time.perf_counter()
            run(C, level_iterations=level_iterations)
            secs.append(time.perf_counter() - t0)
            Fc = np.asarray(C.deformedVoxelPositions, np.float64)
        return float(np.median(secs)), Fc, secs

    sec, Fc, secs = timed_run(counts)
    timed[counts[-1]] = (sec, Fc)
    out["converged_timed_s"] = round(sec, 4)
    out["timed_repeats_s"] = [round(v, 4) for v in secs]
    for name, kk in crit.items():
        if name == "oracle_field" and ref is None:
            continue
        if name == "objective" and ref_objective is None:
            continue
        ent = {"reached": kk is not None}
        if kk is not None:
            if kk not in timed:
                sec, Fc, _ = timed_run(counts[:-1] + [kk])
                timed[kk] = (sec, Fc)
            tt_s, Fc = timed[kk]
            ent.update({"iterations": counts[:-1] + [kk], "gpu_s": round(tt_s, 4),
                        "rms_vox_at_stop_vs_own": round(_rms(Fc, F), 4)})
            if ref is not None:
                ent["rms_vox_at_stop_vs_oracle"] = round(_rms(Fc, ref), 4)
            log(f"to tolerance ({name}): {tt_s:.3f}s ({kk} of {counts[-1]} last-level "
                f"iterations)")
        else:
            log(f"to tolerance ({name}): not reached in {counts[-1]} last-level iterations")
        tol[name] = ent
    out["tolerance_vox_rms"] = TOLERANCE_VOX_RMS
    out["to_tolerance"] = tol
    # the headline criterion: the oracle field when it is reached, else the
    # cold objective (warm start), else the pipeline's own converged field
    for name in ("oracle_field", "objective", "own_field"):
        if tol.get(name, {}).get("reached"):
            out["criterion"] = name
            out["gpu_s"] = tol[name]["gpu_s"]
            out["iterations"] = tol[name]["iterations"]
            break
    k_stop = crit[out["criterion"]] if "criterion" in out else counts[-1]
    return A, F, out, timed.get(k_stop, (None, F))[1]


def diff_breakdown(pair, Fa, Fb):
    """Where two fields differ: RMS (voxels) over tissue and background
    voxels of the fixed image, over edge voxels (gradient magnitude of the
    lightly smoothed fixed image above 0.05 per voxel) and flat tissue, the
    share of the squared difference carried by the worst 1 % of voxels, and
    the maximum."""
    from scipy import ndimage
    d2 = np.sum((np.asarray(Fa, np.float64) - Fb) ** 2, -1)
    fx = ndimage.gaussian_filter(np.squeeze(np.asarray(pair.fixed, np.float64)), 1.0)
    grad = np.sqrt(sum(g ** 2 for g in np.gradient(fx)))
    tissue = (fx > 0.05).reshape(d2.shape)
    edge = (grad > 0.05).reshape(d2.shape)

    def r(m):
        return round(float(np.sqrt(d2[m].mean())), 4) if m.any() else None
    flat = np.sort(d2.ravel())[::-1]
    top = max(1, flat.size // 100)
    return {"rms_all": r(np.ones_like(tissue)), "rms_tissue": r(tissue),
            "rms_background": r(~tissue), "rms_edge": r(edge),
            "rms_flat_tissue": r(tissue & ~edge),
            "voxel_fraction": {"tissue": round(float(tissue.mean()), 4),
                               "edge": round(float(edge.mean()), 4)},
            "top1pct_share_of_squared_diff": round(float(flat[:top].sum() / flat.sum()), 4)
            if flat.sum() > 0 else 0.0,
            "max_vox": round(float(np.sqrt(flat[0])), 4)}


def oracle_pyramid(args, pair, levels=None):
    """The FP64 oracle running the same pipeline as the GPU (pyramid EM to
    its own convergence at every level, 3D topology correction after each),
    on all host cores, timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pycore
    import binder_amd.FastRegistration as frm

    th = host_threads()
    levels = list(levels or args.resolutions)
    saved = frm.Registration
    frm.Registration = pycore
    try:
        O = frm.FastRegistration(voxels=pair.fixed, nodes=pair.moving, threads=th,
                                 **dict(REG_KW, convergence_th=args.convergence_th))
        t0 = time.perf_counter()
        O.run(levels, args.em_iters)
        cpu_s = time.perf_counter() - t0
    finally:
        frm.Registration = saved
    Fo = np.asarray(O.deformedVoxelPositions, np.float64)
    info = {"seconds": round(cpu_s, 3), "cores": th, "kind": "port",
            "iterations": [int(lv[1]) for lv in O.level_log],
            "level_seconds": [round(lv[2], 3) for lv in O.level_log],
            "final_objective": round(float(O.registratorObj.get_objective_trace()[-1, 3]), 2),
            "sample": "FP64 oracle: the whole pyramid to its own convergence with "
                      "topology correction after every level, measured wall-clock"}
    log(f"FP64 oracle pyramid {levels}: iterations {info['iterations']} in {cpu_s:.1f}s "
        f"({th} threads)")
    return Fo, info


def load_warm_net(args):
    from binder_amd import warmstart
    path = args.warm_weights or warmstart.WEIGHTS_C2
    if not os.path.exists(path):
        return None, path, {}
    net = warmstart.WarmStartNet.load(path, max_batch=1)
    with np.load(path, allow_pickle=False) as z:
        meta = {"weights": os.path.relpath(path, ROOT), "params": int(z["params"].size),
                "train_pairs": int(z["train_seeds"].size) if "train_seeds" in z else None,
                "holdout_pairs": int(z["holdout_seeds"].size) if "holdout_seeds" in z else None}
        if "history" in z and z["history"].size:
            h = z["history"]
            meta.update({"epochs": int(h[-1, 0]) + 1, "train_loss": round(float(h[-1, 1]), 6),
                         "holdout_loss": round(float(h[-1, 2]), 6)})
        if "identity_loss" in z:
            meta["holdout_loss_of_identity"] = round(float(z["identity_loss"]), 6)
        meta["holdout_seeds"] = z["holdout_seeds"].tolist() if "holdout_seeds" in z else []
    return net, path, meta


def warm_start_comparison(args, pair, F_cold, cold_out, ref=None):
    """The learned warm start on the same pair: its time to tolerance by
    each criterion (the oracle's cold field, the cold start's final
    objective, its own converged field) against the cold start's, and its
    converged field and objective judged against the cold start's."""
    net, path, meta = load_warm_net(args)
    if net is None:
        log(f"no trained warm-start network at {path}")
        return {"status": f"no trained network at {os.path.relpath(path, ROOT)}"}
    meta.pop("holdout_seeds", None)
    # inference time of the network alone (host bins -> coefficients)
    from binder_amd import warmstart
    f, m = warmstart.pair_inputs(pair.binned_fixed, pair.binned_moving, net.grid)
    net.predict(f, m)
    t0 = time.perf_counter()
    for _ in range(10):
        net.predict(f, m)
    infer_ms = 1e3 * (time.perf_counter() - t0) / 10
    A, Fw, out, _ = registration(args, pair, net=net, ref=ref,
                                 ref_objective=cold_out.get("final_objective"))
    out["network"] = meta
    out["inference_ms"] = round(infer_ms, 3)
    out["rms_vs_cold_converged_vox"] = round(_rms(Fw, F_cold), 4)
    out["start_rms_vs_cold_converged_vox"] = round(
        _rms(warmstart.coefficients_field(A.warm_coef, F_cold.shape[:3]), F_cold), 4)
    # speed-up by criterion: warm time to tolerance against the cold
    # pipeline's time by the same criterion (for the objective criterion the
    # cold pipeline reaches its own final objective at convergence)
    sp = {}
    wt, ct = out.get("to_tolerance", {}), cold_out.get("to_tolerance", {})
    for name in ("oracle_field", "own_field"):
        if wt.get(name, {}).get("reached") and ct.get(name, {}).get("reached"):
            sp[name] = round(ct[name]["gpu_s"] / wt[name]["gpu_s"], 3)
    if wt.get("objective", {}).get("reached") and "converged_timed_s" in cold_out:
        # the cold start reaches its own final objective only at convergence
        sp["objective_vs_cold_converged"] = round(cold_out["converged_timed_s"] /
                                                  wt["objective"]["gpu_s"], 3)
    out["speedup_vs_cold"] = sp
    log(f"warm start: speed-ups {sp}; converged field {out['rms_vs_cold_converged_vox']} vox "
        f"RMS from the cold one, objective {out.get('final_objective_minus_reference')} "
        f"relative to the cold one")
    return out


def warmstart_report(args):
    """Warm against cold start over held-out pairs of the training
    generator (seeds the network never saw)."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    from binder_amd import warmstart_train

    net, path, meta = load_warm_net(args)
    if net is None:
        return {"report": "warmstart", "status": f"no trained network at {path}"}
    first = args.warmstart_first
    seeds = meta.pop("holdout_seeds")[first:first + args.warmstart_eval]
    rows = []
    with ProcessPoolExecutor(max_workers=4, mp_context=mp.get_context("spawn")) as ex:
        for seed, pair in zip(seeds, ex.map(warmstart_train.make_pair, seeds)):
            _, Fc, cold, _ = registration(args, pair)
            _, Fw, warm, _ = registration(args, pair, net=net,
                                          ref_objective=cold["final_objective"])
            wo = warm["to_tolerance"].get("objective", {})
            row = {"seed": int(seed), "cold_s": cold["gpu_s"],
                   "cold_converged_s": cold["converged_timed_s"],
                   "warm_own_s": warm["to_tolerance"]["own_field"].get("gpu_s"),
                   "warm_to_cold_objective_s": wo.get("gpu_s"),
                   "cold_iterations": cold["iterations"],
                   "warm_iterations": warm["converged_iterations"],
                   "warm_used": warm["warm_start_used"],
                   "rms_warm_vs_cold_vox": round(_rms(Fw, Fc), 4),
                   "cold_objective": cold["final_objective"],
                   "warm_objective": warm["final_objective"]}
            rows.append(row)
            log(f"held-out seed {seed}: cold {row['cold_s']:.3f}s, warm to the cold "
                f"objective {row['warm_to_cold_objective_s']}, own {row['warm_own_s']} "
                f"(used {row['warm_used']}), fields {row['rms_warm_vs_cold_vox']} vox apart")
    cs = np.array([r["cold_s"] for r in rows])
    cc = np.array([r["cold_converged_s"] for r in rows])
    ws = np.array([r["warm_own_s"] for r in rows])
    reached = [r for r in rows if r["warm_to_cold_objective_s"] is not None]
    summary = {"pairs": len(rows), "cold_s_median": float(np.median(cs)),
               "warm_own_s_median": float(np.median(ws)),
               "speedup_own_field_per_pair_median": float(np.median(cs / ws)),
               "warm_reaches_cold_objective_fraction": len(reached) / len(rows),
               "speedup_to_cold_objective_per_pair_median": float(np.median(
                   [r["cold_converged_s"] / r["warm_to_cold_objective_s"] for r in reached]))
               if reached else None,
               "cold_converged_s_median": float(np.median(cc)),
               "warm_used_fraction": float(np.mean([r["warm_used"] for r in rows])),
               "rms_warm_vs_cold_median_vox": float(np.median(
                   [r["rms_warm_vs_cold_vox"] for r in rows])),
               "warm_objective_lower_fraction": float(np.mean(
                   [r["warm_objective"] <= r["cold_objective"] for r in rows])),
               "objective_rel_diff_median": float(np.median(
                   [(r["warm_objective"] - r["cold_objective"]) / abs(r["cold_objective"])
                    for r in rows]))}
    return {"report": "warmstart", "network": meta, "warm_levels": list(args.warm_levels),
            "tolerance_vox_rms": TOLERANCE_VOX_RMS, "summary": summary, "pairs": rows}


def batched_report(args):
    """Many pairs and many chains: `--batched N` pairs of uneven shapes
    (axes uniform on {96..192}, warp RMS U(2,6)) registered with several
    pairs in flight, then 16 independent chains per pair of 1e6 site
    updates each from the registered state (graph-captured sweeps).
    Reports pairs/s and chain voxel-updates/s, and the same work one pair
    at a time for comparison."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    from binder_amd import batch

    specs = batch.batch_specs(args.batched, seed=args.seed)
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=16, mp_context=mp.get_context("spawn")) as ex:
        pairs = list(ex.map(batch.make_pair, specs))
    log(f"{len(pairs)} pairs generated in {time.perf_counter() - t0:.1f}s")
    nvox = [int(np.prod(sp["shape"])) for sp in specs]
    out = {"report": "batched", "pairs": len(pairs), "workers": args.batch_workers,
           "shapes": [list(sp["shape"]) for sp in specs],
           "voxels_total": int(sum(nvox))}
    # warm-up: one pair (library set-up, first kernel loads)
    batch.register_pairs(pairs[:1], workers=1)
    regs, dt = batch.register_pairs(pairs, workers=args.batch_workers)
    out["registration"] = {"seconds": round(dt, 3), "pairs_per_s": round(len(pairs) / dt, 2),
                           "iterations": [[int(lv[1]) for lv in r.level_log] for r in regs]}
    log(f"registration: {len(pairs)} pairs in {dt:.2f}s ({len(pairs) / dt:.1f} pairs/s)")
    k = min(len(pairs), args.batch_serial)
    _, dts = batch.register_pairs(pairs[:k], workers=1)
    out["registration"]["one_at_a_time_pairs_per_s"] = round(k / dts, 2)
    # level-synchronous lockstep pyramid with batched launches
    batch.register_pairs_batched(pairs[:2], workers=2)  # warm-up
    lregs, dtl = batch.register_pairs_batched(pairs, workers=args.batch_workers)
    same = all([lv[1] for lv in a.level_log] == [lv[1] for lv in b.level_log] and
               np.array_equal(a.deformedVoxelPositions, b.deformedVoxelPositions)
               for a, b in zip(lregs, regs))
    out["registration_lockstep"] = {"seconds": round(dtl, 3),
                                    "pairs_per_s": round(len(pairs) / dtl, 2),
                                    "identical_to_threaded": bool(same),
                                    "scheduler": batch.register_pairs_batched.stats}
    log(f"lockstep registration: {len(pairs)} pairs in {dtl:.2f}s "
        f"({len(pairs) / dtl:.1f} pairs/s), identical: {same}")
    chains, upd = 16, 1e6
    for use_graph in (True, False):
        n, dtc = batch.chains_for_pairs(regs, chains, upd, workers=args.batch_workers,
                                        use_graph=use_graph)
        key = "chains_graph" if use_graph else "chains_eager"
        out[key] = {"chains_per_pair": chains, "site_updates_per_chain": upd,
                    "voxel_updates": n, "seconds": round(dtc, 3),
                    "voxel_updates_per_s": round(n / dtc, 1)}
        log(f"{key}: {n:.3e} voxel updates in {dtc:.2f}s ({n / dtc:.3e}/s)")
    n1, dt1 = batch.chains_for_pairs(regs[:k], chains, upd, workers=1, use_graph=True)
    out["chains_graph"]["one_pair_at_a_time_voxel_updates_per_s"] = round(n1 / dt1, 1)
    return out


def cnn_report(args):
    """The warm-start network alone at the 160x192x160 workload's input grid
    (80x96x80, two input channels): inference latency of one pair and the
    time of one training step (batch 8: forward, MSE, backward, Adam), with
    the convolutions' algorithmic flop rate (2 * 27 * cin * cout per output
    voxel per layer, x3 for a training step: forward, input and weight
    gradients; the first layer's input gradient is not computed)."""
    from binder_amd import warmstart

    grid = warmstart.network_grid((160, 192, 160))
    rng = np.random.default_rng(args.seed)
    n = 16
    fb = rng.integers(0, 32, (n,) + grid).astype(np.uint8)
    mb = rng.integers(0, 32, (n,) + grid).astype(np.uint8)
    if os.path.exists(warmstart.WEIGHTS_C2):
        net = warmstart.WarmStartNet.load(warmstart.WEIGHTS_C2, max_batch=8)
    else:
        net = warmstart.WarmStartNet(grid, max_batch=8, seed=args.seed)
    C = net.ncoef
    tg = (0.1 * rng.standard_normal((n, 3, C, C, C))).astype(np.float32)
    net.set_dataset(fb, mb, tg)
    nv = [int(np.prod([g >> lvl for g in grid])) for lvl in range(4)]
    levels = [0, 0, 1, 1, 2, 2, 3, 3, 3, 3, 2, 2]
    fwd = sum(2.0 * 27 * ci * co * nv[lv]
              for (ci, co), lv in zip(warmstart.CONV_LAYERS, levels))
    bwd_data = sum(2.0 * 27 * ci * co * nv[lv]
                   for i, ((ci, co), lv) in enumerate(zip(warmstart.CONV_LAYERS, levels)) if i)
    for _ in range(3):
        net.predict(fb[0], mb[0])
    reps = max(args.steps // 4, 10)
    t0 = time.perf_counter()
    for _ in range(reps):
        net.predict(fb[0], mb[0])
    infer_s = (time.perf_counter() - t0) / reps
    idx = np.stack([rng.choice(n, 8, replace=False) for _ in range(reps + 3)]).astype(np.int32)
    net.train(idx[:3])
    t0 = time.perf_counter()
    net.train(idx[3:])
    step_s = (time.perf_counter() - t0) / reps
    train_flops = 8 * (2 * fwd + bwd_data)
    return {"report": "cnn", "grid": list(grid), "params": net.nparam,
            "inference": {"batch": 1, "ms": round(1e3 * infer_s, 3),
                          "conv_gflop": round(fwd / 1e9, 2),
                          "conv_tflops": round(fwd / infer_s / 1e12, 2),
                          "note": "wall-clock incl. host bins upload and coefficient download"},
            "training_step": {"batch": 8, "ms": round(1e3 * step_s, 3),
                              "conv_gflop": round(train_flops / 1e9, 2),
                              "conv_tflops": round(train_flops / step_s / 1e12, 2)},
            "bf16_dense_peak_tflops": 2500.0}


def run(args):
    from binder_amd import _core, synth

    # launch shapes tuned on first use per problem size (results unchanged);
    # every timed region below runs after an untimed run of the same sizes
    _core.autotune(0 if args.no_autotune else 1)
    t0 = time.perf_counter()
    pair, prob = synth.config_problem("c2", seed=args.seed)
    Nx, Ny, Nz = prob.shape
    nv = Nx * Ny * Nz
    N = 3
    gen_s = time.perf_counter() - t0
    log(f"inputs {Nx}x{Ny}x{Nz} generated in {gen_s:.1f}s")

    # ---- FP64 oracle pipeline: the tolerance reference and the CPU time ----
    ref_F, orc_info = None, None
    if not args.no_cpu and not args.quick:
        ref_F, orc_info = oracle_pyramid(args, pair)

    # ---- plain pyramid registration and its time to tolerance -------------
    A, F, reg_out, _ = registration(args, pair, ref=ref_F)
    if orc_info is not None:
        reg_out["cpu_oracle"] = orc_info
        reg_out["diff_vs_oracle"] = diff_breakdown(pair, F, ref_F)
        if "gpu_s" in reg_out:
            reg_out["speedup_vs_cpu"] = round(orc_info["seconds"] / reg_out["gpu_s"], 1)
        reg_out["speedup_vs_cpu_converged"] = round(orc_info["seconds"] /
                                                    reg_out["converged_s"], 1)
    t0 = time.perf_counter()
    _, folds = _core.fold_mask_3D(F.ravel(), Nx, Ny, Nz, 1.0, 1.0, 1.0, e1=0.01)
    fold_s = time.perf_counter() - t0
    reg_out["folds_after"] = folds
    reg_out["fold_detect_s"] = round(fold_s, 5)
    disp_rms = float(np.sqrt(np.mean(np.sum((F - prob.voxel_pos) ** 2, -1))))
    reg_out["displacement_rms_vox"] = round(disp_rms, 4)

    warm_out = None
    if not args.no_warm and not args.quick:
        warm_out = warm_start_comparison(args, pair, F, reg_out, ref=ref_F)
        # the other first levels after the prediction, in brief
        variants = []
        main_levels = list(args.warm_levels)
        for lv in args.warm_variants:
            levels = [int(v) for v in lv.split(",")]
            if levels == main_levels:
                continue
            args.warm_levels = levels
            w = warm_start_comparison(args, pair, F, reg_out, ref=ref_F)
            variants.append({k: w.get(k) for k in (
                "resolutions", "converged_iterations", "converged_timed_s",
                "converged_rms_vs_oracle_vox", "final_objective_minus_reference",
                "rms_vs_cold_converged_vox", "to_tolerance", "speedup_vs_cold")})
        args.warm_levels = main_levels
        if warm_out is not None and variants:
            warm_out["variants"] = variants

    def full_res_handle():
        h = A._new_registration(np.ascontiguousarray(F))
        h.setMILikelihood(A.alpha, np.ascontiguousarray(A.theta).ravel(), 32, 32, 1,
                          np.ascontiguousarray(A.binned_nodes).ravel(),
                          np.ascontiguousarray(A.binned_voxels).ravel())
        return h

    # ---- Gibbs sweeps from the registration result --------------------------
    s = full_res_handle()
    s.sampler_run(args.seed, 0, max(args.warmup, 0) or 0)  # untimed warmup sweeps
    t0 = time.perf_counter()
    s.sampler_run(args.seed + 1, 0, args.steps)
    sweep_s = time.perf_counter() - t0
    ms_per_step = 1e3 * sweep_s / args.steps
    value = nv * args.steps / sweep_s
    log(f"sweeps: {args.steps} in {sweep_s:.3f}s")

    # ---- per-kernel times from HIP events on the handle's stream -----------
    s.set_profiling(True)
    prof_sweeps = WELFORD_BATCH  # one full accumulator batch
    s.sampler_run(args.seed + 2, 0, prof_sweeps)
    ktimes = s.get_kernel_times()
    s.set_profiling(False)
    kernels = {}
    for name, (ms, cnt) in ktimes.items():
        avg_us = 1e3 * ms / max(cnt, 1)
        b, fl = kernel_model(name, nv, nv, N, (Nx, Ny, Nz))
        ent = {"avg_us": round(avg_us, 2), "launches_per_sweep": cnt / prof_sweeps,
               "ms_per_sweep": round(ms / prof_sweeps, 4)}
        if b is not None:
            ent["GBps"] = round(b / (avg_us * 1e-6) / 1e9, 1)
            ent["hbm_frac"] = round(ent["GBps"] / HBM_PEAK_GBS, 4)
            if fl:
                ent["TFLOPs"] = round(fl / (avg_us * 1e-6) / 1e12, 2)
                ent["bf16_mfma_frac"] = round(ent["TFLOPs"] / BF16_PEAK_TFLOPS, 4)
        kernels[name] = ent
    modelled = [k for k in kernels if kernel_model(k, nv, nv, N, (Nx, Ny, Nz))[0] is not None]
    dom = max(modelled, key=lambda k: kernels[k]["ms_per_sweep"])
    b, fl = kernel_model(dom, nv, nv, N, (Nx, Ny, Nz))
    avg_s = kernels[dom]["avg_us"] * 1e-6
    if fl:
        roof = {"kernel": dom, "bound": "mfma", "achieved": round(fl / avg_s / 1e12, 3),
                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(b / avg_s / 1e9, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["traffic"] = MEASURED_TRAFFIC.get((Nx, Ny, Nz), {}).get(dom)
    roof["bytes_per_launch"] = b
    roof["bytes_model"] = "32 B/voxel: f32 field in, f32 dream out, int32 fixed and moving bins"
    if dom in ("estep", "estep_sample"):
        pb = packed_bin_bytes(nv, nv, N)
        roof["frac_u8_bins"] = round(pb / avg_s / 1e9 / HBM_PEAK_GBS, 4)
    # the whole sweep against its own minimum traffic (264 B per site update)
    sweep_gbs = SWEEP_BYTES_PER_UPDATE * value / 1e9
    roof["sweep"] = {"bytes_per_update": SWEEP_BYTES_PER_UPDATE,
                     "achieved": round(sweep_gbs, 1), "unit": "GB/s",
                     "frac": round(sweep_gbs / HBM_PEAK_GBS, 4)}

    # EM iteration at full resolution: wall-clock and kernel times
    reg = full_res_handle()
    reg.EM(2, -1e30, 1, 0)
    t0 = time.perf_counter()
    reg.EM(20, -1e30, 1, 0)
    em_it_ms = 1e3 * (time.perf_counter() - t0) / 20
    reg.set_profiling(True)
    reg.EM(2, -1e30, 1, 0)
    em_k = {k: round(1e3 * v[0] / max(v[1], 1), 2) for k, v in reg.get_kernel_times().items()}
    reg.set_profiling(False)

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "Gibbs sweep, 3D MI registration 160x192x160, 32 bins, "
                               "after cold-start pyramid EM + topology correction",
                   "shape": [Nx, Ny, Nz], "bins": prob.K, "gamma_per_voxel": 1e-6,
                   "alpha": prob.alpha, "seed": args.seed},
        "roofline": roof,
        "registration": reg_out,
        "warm_start": warm_out,
        "em_full_res": {"ms_per_iteration": round(em_it_ms, 4), "kernel_avg_us": em_k},
        "kernels": kernels,
        "input_gen_s": round(gen_s, 2),
        "autotune": [{"op": r[0], "key": r[1], "choice": r[2], "default": r[3],
                      "ms_choice": r[4], "ms_default": r[5]} for r in _core.autotune_report()],
    }
    if not args.no_cpu:
        log("CPU baseline (oracle, all host cores)")
        out["cpu_baseline"] = cpu_baseline(prob, nv)
    return out


def cpu_baseline(prob, nv):
    """FP64 CPU oracle (restatement of the reference CPU backend) on the same
    workload: one Gibbs sweep over the full volume with all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pycore
    from binder_amd import synth

    th = host_threads()
    r = synth.build_registration(pycore, prob, threads=th)
    t0 = time.perf_counter()
    r.sampler(1, 0, 1, 0, 1, 1, 1, 1)  # probe: one sweep
    t1 = time.perf_counter() - t0
    sweeps = int(max(2, min(30, round(10.0 / max(t1, 1e-3)))))
    t0 = time.perf_counter()
    r.sampler(2, 0, sweeps, 0, 1, 1, 1, 1)
    t = time.perf_counter() - t0
    return {"value": round(nv * sweeps / t, 1), "unit": "voxel-updates/s", "cores": th,
            "kind": "port", "sample": f"{sweeps} Gibbs sweeps of the FP64 oracle over the full "
                                      "160x192x160 volume (E-draw, Dirichlet, DCT-domain draw, "
                                      f"Welford) after a 1-sweep probe, {t:.2f} s"}


def _dct_matrix(n):
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    C = np.sqrt(2.0 / n) * np.cos(np.pi * (2 * j + 1) * k / (2 * n))
    C[0] /= np.sqrt(2.0)
    return C  # C[k][j], orthonormal DCT-II


def _mi(H):
    p = H / H.sum()
    pk = p.sum(1, keepdims=True)
    pl = p.sum(0, keepdims=True)
    nz = p > 0
    return float(np.sum(p[nz] * np.log(p[nz] / (pk @ pl)[nz])))


def c1_report(args):
    """The 2D configuration (192x160, 32 bins): the FP64 oracle's pyramid
    to its own convergence (all host cores, timed), the GPU pyramid's time
    to within 0.1 px RMS of that field, where the two differ, and the
    oracle's own sensitivity to a 1e-4 px start perturbation."""
    from binder_amd import synth
    import binder_amd.FastRegistration as frm
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pycore

    pair, _ = synth.config_problem("c1", seed=args.seed)
    ref_F, orc = oracle_pyramid(args, pair)
    A, F, reg, _ = registration(args, pair, ref=ref_F)
    reg["cpu_oracle"] = orc
    if "gpu_s" in reg:
        reg["speedup_vs_cpu"] = round(orc["seconds"] / reg["gpu_s"], 1)
    reg["speedup_vs_cpu_converged"] = round(orc["seconds"] / reg["converged_s"], 1)
    reg["diff_vs_oracle"] = diff_breakdown(pair, F, ref_F)
    # conditioning of the pipeline itself: the FP64 oracle against itself
    # with the initial positions moved by N(0, 1e-4) px
    th = host_threads()
    saved = frm.Registration
    frm.Registration = pycore
    try:
        P = frm.FastRegistration(voxels=pair.fixed, nodes=pair.moving, threads=th,
                                 **dict(REG_KW, convergence_th=args.convergence_th))
        P.full_voxels_pos = P.full_voxels_pos + np.random.default_rng(1).normal(
            0.0, 1e-4, P.full_voxels_pos.shape)
        P.deformedVoxelPositions = P.full_voxels_pos
        P.run(list(args.resolutions), args.em_iters)
    finally:
        frm.Registration = saved
    reg["oracle_self_rms_vox_at_1e-4_start_perturbation"] = round(
        _rms(P.deformedVoxelPositions, ref_F), 5)
    reg["oracle_self_iterations_at_1e-4_start_perturbation"] = [
        int(lv[1]) for lv in P.level_log]
    return {"report": "c1", "config": {"shape": list(pair.fixed.shape), "bins": 32,
                                       "gamma_per_voxel": 1e-6, "seed": args.seed},
            "registration": reg}


def precision_report(args):
    """Error of the GPU path against the FP64 oracle at 256x256x192 (the
    precision variant): field, warped image, MI and objective after the same
    EM iterations from the same start; fold counts on the same field; the
    sampler's voxel-wise variance against its closed form."""
    from binder_amd import _core, resample, synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pycore

    th = host_threads()
    t0 = time.perf_counter()
    pair, prob = synth.config_problem("c3", seed=args.seed)
    Nx, Ny, Nz = prob.shape
    log(f"c3 inputs {prob.shape} in {time.perf_counter() - t0:.1f}s")
    iters = args.precision_iters
    g = synth.build_registration(_core, prob)
    o = synth.build_registration(pycore, prob, threads=th)
    g.EM(iters, -1e30, 1, 0)
    t0 = time.perf_counter()
    o.EM(iters, -1e30, 1, 0)
    log(f"oracle EM {iters} iterations in {time.perf_counter() - t0:.1f}s")
    fg = g.get_final_locations_as_numpy_array().astype(np.float64)
    fo = o.get_final_locations_as_numpy_array()
    d = fg - fo
    field = {"rms_vox": float(np.sqrt(np.mean(np.sum(d ** 2, -1)))),
             "max_vox": float(np.abs(d).max())}
    tg, to = g.get_objective_trace(), o.get_objective_trace()
    objective = {"mll_rel": float(abs(tg[-1, 0] - to[-1, 0]) / abs(to[-1, 0])),
                 "prior_d_rel": float(abs(tg[-1, 2] - to[-1, 2]) / max(abs(to[-1, 2]), 1e-300)),
                 "mlp_rel": float(abs(tg[-1, 3] - to[-1, 3]) / abs(to[-1, 3]))}
    _, Hg = g.e_step(return_histogram=True)
    _, Ho = o.e_step(return_histogram=True)
    mi = {"gpu": _mi(Hg), "oracle": _mi(Ho)}
    mi["rel"] = abs(mi["gpu"] - mi["oracle"]) / abs(mi["oracle"])
    wg = resample.warp(pair.moving, np.moveaxis(fg, -1, 0), order=1)
    wo = resample.warp(pair.moving, np.moveaxis(fo, -1, 0), order=1)
    rng_int = float(pair.moving.max() - pair.moving.min())
    warped = {"rms_rel_range": float(np.sqrt(np.mean((wg - wo) ** 2)) / rng_int),
              "max_rel_range": float(np.abs(wg - wo).max() / rng_int)}
    # folds: GPU and oracle Jacobians of the same (GPU) field
    _, cg = _core.fold_mask_3D(fg.ravel(), Nx, Ny, Nz, 1.0, 1.0, 1.0, e1=0.0)
    Jo = pycore.compute_jacobian_determinant_3D(fg.ravel(), Nx, Ny, Nz, 1.0, 1.0, 1.0)
    Jg = _core.compute_jacobian_determinant_3D(fg.ravel(), Nx, Ny, Nz, 1.0, 1.0, 1.0)
    folds = {"gpu": int(cg), "oracle": int(np.count_nonzero(Jo < 0.0)),
             "jacobian_bitexact": bool(np.array_equal(Jg, Jo))}
    # covariance: deformation block only, dream = x, so each component is
    # x + DCT-III(sqrt(sigma^2 / (1 + gamma sigma^2 D)) eps): voxel variance
    # sum_k C[k, i]^2 sigma^2 / (1 + gamma sigma^2 D_k), separable
    sig2 = 9.0 / (8.0 * np.pi)
    gam = prob.gammas[0]
    V = sig2 / (1.0 + gam * sig2 * prob.D)
    var = V
    for ax, n in enumerate(prob.shape):
        C2 = _dct_matrix(n) ** 2
        var = np.moveaxis(np.tensordot(C2.T, np.moveaxis(var, ax, 0), axes=1), 0, ax)
    s = synth.build_registration(_core, prob)
    s.set_dream_locations(prob.voxel_pos.ravel())
    ns = args.precision_sweeps
    _, cov = s.sampler(args.seed, 0, ns, 0, 1, 0, 0, 1)
    vd = np.stack([cov[..., k, k] for k in range(3)], -1)
    ratio = vd / var[..., None]
    se = np.sqrt(2.0 / (ns - 1))  # relative standard error of a Gaussian variance
    covariance = {"sweeps": ns, "mean_ratio": float(ratio.mean()),
                  "rms_rel_err": float(np.sqrt(np.mean((ratio - 1) ** 2))),
                  "mc_rel_se": float(se),
                  "frac_beyond_4se": float(np.mean(np.abs(ratio - 1) > 4 * se))}
    return {"report": "precision", "config": {"shape": [Nx, Ny, Nz], "bins": prob.K,
                                              "em_iterations": iters,
                                              "gamma_per_voxel": 1e-6, "seed": args.seed},
            "dct": "split-BF16 MFMA (3 products, FP32 accumulation); FP64 cost, Welford, "
                   "Jacobians",
            "field": field, "objective": objective, "mi": mi, "warped_image": warped,
            "folds": folds, "covariance": covariance, "oracle_threads": th}


def copy_calibration(mib=1024):
    """A plain device-to-device copy of known size (torch elementwise copy),
    used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 copies of {mib} MiB (read {mib} MiB + write {mib} MiB each)")


def models_report(args):
    """The reference's other models at 160x192x160, 32 bins (SSD likelihood,
    affine and rigid transformations; BINDER/src/cpu/SSDLikelihood.cpp,
    AffineTransformation.cpp, RigidTransformation.cpp) beside MI + DCT prior:
    per EM iteration and per Gibbs sweep on the GPU, and 3 EM iterations from
    the identity against the FP64 oracle (all host cores)."""
    from binder_amd import _core as core, synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pycore as orc
    core.autotune(0 if args.no_autotune else 1)
    pair, prob = synth.config_problem("c2", seed=args.seed)
    nv = int(np.prod(prob.shape))
    sig0 = float(np.mean((prob.binned_nodes.astype(np.float64) - prob.binned_voxels) ** 2))
    th = host_threads()

    def make(be, kind, threads=1):
        Nx, Ny, Nz = prob.shape
        r = be.Registration(prob.final_locations.ravel(), threads, prob.voxel_pos.ravel(),
                            prob.voxel_pos.ravel(), Nx, Ny, Nz, Nx, Ny, Nz, 3, prob.nnz)
        r.setNonLinearTransformation(*prob.gammas, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, prob.D.ravel())
        if kind == "affine":
            r.setAffineTransformation(np.eye(4).ravel(), np.zeros(3))
        if kind == "rigid":
            r.setRigidTransformation(np.eye(3).ravel(), np.zeros(3))
        if kind == "ssd":
            r.setSSDLikelihood(prob.K, threads, sig0, 1, 1.0, 0.0, prob.binned_nodes.ravel(),
                               prob.binned_voxels.ravel(), 1.0, 1.0)
        else:
            r.setMILikelihood(prob.alpha, prob.theta.ravel(), prob.K, prob.L, threads,
                              prob.binned_nodes.ravel(), prob.binned_voxels.ravel())
        return r

    out = {"workload": "160x192x160, 32 bins, c2 synthetic pair (seed %d)" % args.seed,
           "models": {}}
    for kind in ("mi", "ssd", "affine", "rigid"):
        ent = {}
        g = make(core, kind)
        g.EM(3, -1e30, 1, 0)
        fg = np.asarray(g.get_final_locations_as_numpy_array(), np.float64)
        tg = g.get_objective_trace()
        t0 = time.perf_counter()
        o = make(orc, kind, th)
        o.EM(3, -1e30, 1, 0)
        ent["oracle_3_iterations_s"] = round(time.perf_counter() - t0, 2)
        fo = o.get_final_locations_as_numpy_array()
        to = o.get_objective_trace()
        ent["vs_oracle_3_iterations"] = {
            "field_max_vox": float(np.abs(fg - fo).max()),
            "mll_rel": float(np.max(np.abs(tg[:, 0] - to[:, 0]) / np.abs(to[:, 0]))),
            "mlp_rel": float(np.max(np.abs(tg[:, 3] - to[:, 3]) / np.abs(to[:, 3])))}
        if kind == "ssd":
            ent["ssd_parameters_gpu_oracle"] = [g.get_ssd_parameters(), o.get_ssd_parameters()]
        if kind in ("affine", "rigid"):
            pg, po = np.array(g.get_linear_parameters()), np.array(o.get_linear_parameters())
            ent["linear_parameters_max_diff"] = float(np.abs(pg - po).max())
        del o
        g.EM(3, -1e30, 1, 0)  # warm
        t0 = time.perf_counter()
        g.EM(20, -1e30, 1, 0)
        ent["em_ms_per_iteration"] = round(1e3 * (time.perf_counter() - t0) / 20, 4)
        g.set_profiling(True)
        g.EM(5, -1e30, 1, 0)
        ent["em_kernel_avg_us"] = {k: round(1e3 * v[0] / max(v[1], 1), 2)
                                   for k, v in g.get_kernel_times().items()}
        g.set_profiling(False)
        g.sampler_run(args.seed, 2, 8)  # warm
        t0 = time.perf_counter()
        g.sampler_run(args.seed + 1, 0, 40)
        sw = (time.perf_counter() - t0) / 40
        ent["sweep_ms"] = round(1e3 * sw, 4)
        ent["voxel_updates_per_s"] = round(nv / sw, 1)
        out["models"][kind] = ent
        log(f"{kind}: EM {ent['em_ms_per_iteration']} ms/iteration, sweep {ent['sweep_ms']} ms, "
            f"vs oracle {ent['vs_oracle_3_iterations']}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--em-iters", type=int, default=1000,
                    help="EM iterations per level, times the level's resolution "
                         "(run_BINDER --iterations)")
    ap.add_argument("--resolutions", type=int, nargs="+", default=[8, 4, 2, 1])
    ap.add_argument("--convergence-th", type=float, default=2.5e-5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--quick", action="store_true",
                    help="3 EM iterations and no topology iterations (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the copy kernel used to calibrate HBM counters")
    ap.add_argument("--precision", action="store_true",
                    help="only the 256x256x192 precision report against the FP64 oracle")
    ap.add_argument("--precision-iters", type=int, default=3)
    ap.add_argument("--c1", action="store_true",
                    help="only the 2D 192x160 registration report (GPU vs FP64 oracle)")
    ap.add_argument("--precision-sweeps", type=int, default=200)
    ap.add_argument("--warm-variants", nargs="*", default=["1", "4,2,1"],
                    help="further first levels after the prediction, reported in brief "
                         "(comma-separated pyramid levels each)")
    ap.add_argument("--warm-levels", type=int, nargs="+", default=[2, 1],
                    help="pyramid levels run from the learned warm start")
    ap.add_argument("--warm-weights", default=None,
                    help="trained warm-start network (default: the packaged weights)")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start comparison")
    ap.add_argument("--no-autotune", action="store_true",
                    help="default launch shapes instead of autotuned ones")
    ap.add_argument("--warmstart-eval", type=int, default=0, metavar="N",
                    help="only the warm-vs-cold report over N held-out pairs")
    ap.add_argument("--cnn", action="store_true",
                    help="only the warm-start network's inference / training-step report")
    ap.add_argument("--warmstart-first", type=int, default=0, metavar="I",
                    help="first held-out pair of --warmstart-eval (0-based)")
    ap.add_argument("--models", action="store_true",
                    help="only the SSD / affine / rigid model report (GPU times, oracle parity)")
    ap.add_argument("--batched", type=int, default=0, metavar="N",
                    help="only the many-pairs / many-chains report over N pairs")
    ap.add_argument("--batch-workers", type=int, default=8)
    ap.add_argument("--batch-serial", type=int, default=8,
                    help="pairs of the one-at-a-time comparison")
    args = ap.parse_args()
    if args.precision:
        print(json.dumps(precision_report(args)), flush=True)
        return
    if args.c1:
        print(json.dumps(c1_report(args)), flush=True)
        return
    if args.warmstart_eval:
        print(json.dumps(warmstart_report(args)), flush=True)
        return
    if args.batched:
        print(json.dumps(batched_report(args)), flush=True)
        return
    if args.models:
        print(json.dumps(models_report(args)), flush=True)
        return
    if args.cnn:
        print(json.dumps(cnn_report(args)), flush=True)
        return
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the registration and sampling path runs on one GPU; sharding it is a design, not built)", file=sys.stderr)
        sys.exit(2)
    print(json.dumps(run(args)), flush=True)


if __name__ == "__main__":
    main()
