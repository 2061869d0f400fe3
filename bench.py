#!/usr/bin/env python3
"""Force-evaluation throughput of the MI355X gravity engine.

One step = one complete force evaluation of the workload: cell keys, stable
radix sort, cell table, PP pair force and the whole PM chain (deposit, r2c,
Green's function, c2r, gradient + interpolation).  Particles are resident in
HBM before the timed region.  The headline value is the FP32 path's particle
force evaluations per second at 256^3 particles / 512^3 mesh (the clustered
500 Mpc/h box with 2000 NFW halos); the FP64 path, the FP32-vs-FP64 error
percentiles, the PP kernel's roofline position and a CPU baseline (the FP64
oracle on this host) are reported beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle force evals/s at 256^3 (FP32 vs FP64 GPU) + 99th-pct force err"
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector, AMD published
PEAK_HBM_GBS = 8000.0
FLOPS_PER_INTERACTION = 30


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentiles(err):
    return {"median": float(np.median(err)), "p99": float(np.percentile(err, 99)),
            "max": float(err.max())}


def cpu_baseline(box, pos, mass, n_targets, seed=5006, gpu_fp64=None):
    """FP64 oracle on the host: PP on a random target sample (the cell list
    over all particles is built as in a full evaluation) plus the full PM;
    returns the extrapolated whole-box force evals/s.  With gpu_fp64 = the
    GPU FP64 path's (a_pp, a_pm), the oracle's results also check it: PP on
    the sampled targets bit for bit, PM over all particles relative to the
    GPU FP64 |a_total| (whose PP part is the one checked bit for bit)."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = min(threads, os.cpu_count() or threads)
    from cuco_amd import ics
    p = oracle.params(box.M, box.L, ics.G_INTERNAL, box.asmth, box.rcut,
                      box.soft_h, nthreads=threads)
    rng = np.random.default_rng(seed)
    targets = np.sort(rng.choice(box.N, n_targets, replace=False))
    orc_pp, _ = oracle.pp_force(p, pos, mass, targets)
    t_setup, t_sample = oracle.pp_last_timing()
    t_pp = t_setup + t_sample * box.N / n_targets
    t0 = time.perf_counter()
    orc_pm = oracle.pm_force(p, pos, mass)
    t_pm = time.perf_counter() - t0
    parity = None
    if gpu_fp64 is not None:
        g_pp, g_pm = gpu_fp64
        den = np.linalg.norm(g_pp + g_pm, axis=1)
        den = np.where(den > 0, den, 1.0)
        parity = {
            "targets": int(n_targets),
            "pp_bit_identical": bool(np.array_equal(g_pp[targets], orc_pp)),
            "pp_max_rel": float((np.linalg.norm(g_pp[targets] - orc_pp, axis=1) /
                                 den[targets]).max()),
            "pm_max_rel": float((np.linalg.norm(g_pm - orc_pm, axis=1) / den).max()),
        }
        del den
    value = box.N / (t_pp + t_pm)
    # The faithful single-threaded structure: the same PP loop on one core.
    p1 = oracle.params(box.M, box.L, ics.G_INTERNAL, box.asmth, box.rcut,
                       box.soft_h, nthreads=1)
    n1 = max(1000, n_targets // 20)
    oracle.pp_force(p1, pos, mass, targets[:n1])
    _, t_sample1 = oracle.pp_last_timing()
    return parity, {
        "value": value,
        "unit": "particle force evals/s",
        "cores": threads,
        "kind": "port",
        "single_thread_pp_evals_per_s": n1 / t_sample1,
        "sample": (f"FP64 oracle (OpenMP, {threads} threads): PP for {n_targets} "
                   f"random targets of the {box.N}-particle box, cell list over "
                   f"all particles (setup {t_setup:.2f}s, sample {t_sample:.2f}s, "
                   f"extrapolated PP {t_pp:.1f}s); full PM on the {box.M}^3 "
                   f"mesh {t_pm:.1f}s; one thread: PP for {n1} targets in "
                   f"{t_sample1:.2f}s"),
    }


def ensemble_bench(E, ics, nboxes, steps, tune=True, looped=True):
    """The 64-box ensemble: FP32 KDK steps (each one hipGraph launch) of all
    boxes batched in one handle against the same boxes run one after
    another through single-box handles.  The ensemble handle is autotuned
    on its own particles and the single-box handles take the same settings.
    Both sides replay an already captured graph from the same initial
    state; the final states must be bit-identical."""
    boxes = ics.ensemble_boxes(nboxes)
    t0 = time.perf_counter()
    data = [ics.generate(b) for b in boxes]
    log(f"[bench] ensemble: {len(boxes)} boxes, {sum(b.N for b in boxes)} particles "
        f"generated in {time.perf_counter() - t0:.1f}s")
    b0 = boxes[0]
    d, k = E.init_drift_table(b0.a_start, b0.a_end, ics.OMEGA_M, ics.OMEGA_L,
                              ics.HUBBLE_INTERNAL)
    f = E.kdk_factors(d, k, b0.a_start, b0.a_end, b0.a_start, b0.a_end, steps)
    pos = np.concatenate([x[0] for x in data])
    vel = np.concatenate([x[1] for x in data])
    mass = np.concatenate([x[2] for x in data])

    ens = E.Ensemble([b.engine_kwargs() for b in boxes], [b.N for b in boxes])
    tuning = None
    if tune:
        ens.upload(pos, vel, mass)
        t0 = time.perf_counter()
        tuning, _ = ens.autotune()
        log(f"[bench] ensemble autotune {tuning} ({time.perf_counter() - t0:.1f}s)")
    else:
        tuning = ens.tuning()

    # FP32 against the FP64 path on the initial state (every box; the
    # denominator is each particle's FP64 total acceleration).
    ens.upload(pos, vel, mass)
    ens.compute_accelerations(E.FP64, True)
    ref_pp, ref_pm = ens.accelerations()
    ens.compute_accelerations(E.FP32, True)
    a_pp, a_pm = ens.accelerations()
    den = np.linalg.norm(ref_pp + ref_pm, axis=1)
    den = np.where(den > 0, den, 1.0)
    errors = {
        "pp": percentiles(np.linalg.norm(a_pp - ref_pp, axis=1) / den),
        "pm": percentiles(np.linalg.norm(a_pm - ref_pm, axis=1) / den),
        "total": percentiles(np.linalg.norm(a_pp + a_pm - ref_pp - ref_pm, axis=1) / den),
    }
    del ref_pp, ref_pm, a_pp, a_pm, den
    log(f"[bench] ensemble FP32 vs FP64 {errors['total']}")

    def run_batched():
        ens.upload(pos, vel, mass)
        ens.compute_accelerations(E.FP32, True)
        ens.synchronize()
        t = time.perf_counter()
        ens.kdk_steps(f, E.FP32, use_graph=True)
        return time.perf_counter() - t

    run_batched()                       # captures the step graph
    t_batched = run_batched()
    bpos, _ = ens.download_particles()
    # Stage times of eager FP32 evaluations on the final state (outside the
    # timed region; the first evaluation re-sorts after the last drift).
    ens.compute_accelerations(E.FP32, True)
    ens.synchronize()
    ens.set_timing(True)
    ens.stage_times(reset=True)
    for _ in range(3):
        ens.compute_accelerations(E.FP32, True)
    ens.synchronize()
    st, nev = ens.stage_times(reset=True)
    ens.set_timing(False)
    stage_ms = {k: v / max(nev, 1) for k, v in st.items()}
    ens.close()
    log(f"[bench] ensemble batched {t_batched * 1e3 / steps:.2f} ms/step")

    t_looped, identical = 0.0, True
    for b, (box, (p, v, m)) in enumerate(zip(boxes, data) if looped else ()):
        eng = E.Engine(max_particles=box.N, **box.engine_kwargs())
        eng.set_tuning(**tuning)
        for rep in range(2):            # the first run captures the graph
            eng.upload(p, v, m)
            eng.compute_accelerations(E.FP32, True)
            eng.synchronize()
            t = time.perf_counter()
            eng.kdk_steps(f, E.FP32, use_graph=True)
            dt = time.perf_counter() - t
        t_looped += dt
        sp, _ = eng.download_particles()
        o = sum(x.N for x in boxes[:b])
        identical = identical and bool(np.array_equal(bpos[o:o + box.N], sp))
        eng.close()
    if looped:
        log(f"[bench] ensemble looped {t_looped * 1e3 / steps:.2f} ms/step, "
            f"identical={identical}")
    n_total = sum(b.N for b in boxes)
    sides = sorted({b.n_side for b in boxes})
    return {
        "workload": (f"{len(boxes)} boxes, N_b in {{{', '.join(f'{s}^3' for s in sides)}}} "
                     f"(seed 4004), mesh 2 N_b^(1/3), L_b = N_b^(1/3) x 3.125 Mpc/h, "
                     f"Zel'dovich ICs; FP32 KDK steps, one hipGraph per step"),
        "n_boxes": len(boxes),
        "n_particles": n_total,
        "tuning": tuning,
        "steps": steps,
        "batched_ms_per_step": t_batched * 1e3 / steps,
        "looped_ms_per_step": t_looped * 1e3 / steps if looped else None,
        "batched_over_looped": t_looped / t_batched if looped else None,
        "batched_evals_per_s": n_total * steps / t_batched,
        "bit_identical_to_looped": identical if looped else None,
        "error_vs_fp64": errors,
        "stage_ms_fp32": stage_ms,
    }


def timeline_bench(E, ics, box, pos, vel, mass, steps):
    """Individual power-of-two timesteps (run.c / timestep.c) on the
    workload, FP32: `steps` passes of run()'s body from a_start, each a
    drift to the next sync point, the active set's forces (PM on PM steps)
    and its kicks and new timesteps.  Reports the per-step wall time and
    the active particles per second."""
    eng = E.Engine(max_particles=box.N, **box.engine_kwargs())
    eng.upload(pos, vel, mass)
    tl = dict(time_begin=box.a_start, time_max=box.a_end, omega0=ics.OMEGA_M,
              omega_lambda=ics.OMEGA_L, hubble=ics.HUBBLE_INTERNAL,
              err_tol_int_accuracy=0.025,
              max_size_timestep=math.log(box.a_end / box.a_start) / 4,
              min_size_timestep=0.0, softening_halo=box.softening,
              softening_halo_max_phys=box.softening)
    eng.timeline_init(**tl)
    eng.timeline_run(E.FP32, max_steps=1)   # the first step: every particle
    eng.synchronize()
    active, done = 0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        st = eng.timeline_run(E.FP32, max_steps=1)
        active += st["active"]
        done += 1
        if st["finished"]:
            break
    eng.synchronize()
    t = time.perf_counter() - t0
    b, e = eng.timesteps()
    levels = np.unique(e - b)
    eng.close()
    log(f"[bench] timeline: {done} steps, {active / max(done, 1) / box.N:.3f} of the "
        f"particles active per step, {t / max(done, 1) * 1e3:.2f} ms/step")
    return {
        "workload": (f"{box.name}, individual power-of-two timesteps from a = {box.a_start} "
                     f"(ErrTolIntAccuracy 0.025, MaxSizeTimestep ln(a_end/a_start)/4), "
                     f"FP32; {done} steps after the first (all-active) one"),
        "steps": done,
        "ms_per_step": t / max(done, 1) * 1e3,
        "active_per_step": active / max(done, 1),
        "active_evals_per_s": active / t,
        "timestep_levels": [int(x) for x in levels],
    }


def stage_bytes(N: int, M: int, key_bits: int, nkeys: int) -> dict:
    """Compulsory HBM bytes of one FP32 evaluation's stages (every input
    read once, every output written once) for N particles and an M^3 mesh:
    R = M^3 4-byte reals, C = M^2 (M/2+1) 8-byte complex values."""
    R = 4 * M ** 3
    C = 8 * M * M * (M // 2 + 1)
    passes = -(-key_bits // 8)
    return {
        # positions (double4) in, 4-byte Morton key out
        "keys": 36 * N,
        # per 8-bit pass: key read for the digit counts, key/value read and
        # written by the scatter
        "sort": passes * 20 * N,
        # CSR cell offsets (keys in, one entry per key written) + reorder
        # (perm and key in, double4 gathered, cell-relative float4 out)
        "cells": 4 * N + 4 * (nkeys + 1) + 56 * N,
        # sorted float4 + key in, density mesh out
        "deposit": 20 * N + R,
        # z rows r2c (R in, C out) + y lines (C in, C out)
        "fft_fwd": R + 3 * C,
        # x lines: forward FFT, Green's function, inverse FFT in one pass
        "green_x": 2 * C,
        # y lines + z rows c2r
        "fft_inv": 3 * C + R,
        # potential mesh in, sorted float4 + key + perm in, float4 out
        "interp": R + 40 * N,
    }


def hbm_roofline(stage_ms: dict, nbytes: dict, stages) -> dict:
    """Achieved GB/s and the fraction of the 8 TB/s HBM peak per stage and
    summed over `stages`, from HIP-event stage times."""
    out = {}
    tb, tm = 0, 0.0
    for k in stages:
        ms = stage_ms.get(k, 0.0)
        if ms <= 0:
            continue
        gbs = nbytes[k] / (ms * 1e-3) / 1e9
        out[k] = {"bytes": nbytes[k], "ms": ms, "achieved": gbs,
                  "frac": gbs / PEAK_HBM_GBS}
        tb += nbytes[k]
        tm += ms
    if tm > 0:
        out["total"] = {"bytes": tb, "ms": tm, "achieved": tb / (tm * 1e-3) / 1e9,
                        "frac": tb / (tm * 1e-3) / 1e9 / PEAK_HBM_GBS}
    return out


def morton_bits(M: int) -> int:
    return 3 * max(1, math.ceil(math.log2(M)))


def pm_and_sort_roofline(E, eng, box, stage_fp32, reps=3) -> dict:
    """HBM roofline of the sort and PM stages of the FP32 path.  Sort stages
    (keys, sort, cell table) from the timed evaluations, which run them on
    the main stream before anything else; PM stages from PM-only
    evaluations (nothing runs beside them)."""
    eng.set_timing(True)
    eng.pmforce_periodic(E.FP32)
    eng.synchronize()
    eng.stage_times(reset=True)
    for _ in range(reps):
        eng.pmforce_periodic(E.FP32)
    eng.synchronize()
    st, n = eng.stage_times(reset=True)
    eng.set_timing(False)
    pm_ms = {k: v / max(n, 1) for k, v in st.items()}
    kb = morton_bits(box.M)
    nbytes = stage_bytes(box.N, box.M, kb, 1 << kb)
    return {
        "bound": "hbm",
        "unit": "GB/s",
        "peak": PEAK_HBM_GBS,
        "note": ("compulsory bytes per stage (each input read once, each output "
                 "written once) / HIP-event stage time; PM stages timed in "
                 "PM-only evaluations"),
        "sort": hbm_roofline(stage_fp32, nbytes, ["keys", "sort", "cells"]),
        "pm": hbm_roofline(pm_ms, nbytes, ["deposit", "fft_fwd", "green_x", "fft_inv",
                                           "interp"]),
    }


def halo512_bench(E, ics, tuning, steps=5, n_oracle=1_000_000, seed=5006,
                  with_oracle=True):
    """The large single box: 512^3 particles, 1024^3 FP32 mesh, 1000 Mpc/h,
    20000 halos.  `steps` FP32 KDK steps (one hipGraph per step, after a
    warm-up run that captures it), the pair kernel's interactions/s against
    the FP32 vector peak, the PM chain's HBM GB/s, and the FP32 force error
    at the initial conditions against the FP64 oracle on a random subset of
    n_oracle particles (PP for those targets over all particles, PM from
    the oracle's full FP64 mesh)."""
    box = ics.WORKLOADS["halo512"]
    t0 = time.perf_counter()
    pos, vel, mass = ics.generate(box)
    log(f"[bench] halo512: N={box.N} generated in {time.perf_counter() - t0:.1f}s")
    eng = E.Engine(max_particles=box.N, **box.engine_kwargs())
    eng.upload(pos, vel, mass)
    eng.set_tuning(**tuning)
    eng.compute_accelerations(E.FP32, True)
    eng.synchronize()
    a32, p32 = eng.accelerations()
    pairs_in, candidates = eng.count_pairs()
    eng.set_timing(True)
    eng.gravity_linklist(E.FP32)
    eng.synchronize()
    eng.stage_times(reset=True)
    for _ in range(2):
        eng.gravity_linklist(E.FP32)
    eng.synchronize()
    st_pp, n_pp = eng.stage_times(reset=True)
    eng.set_timing(False)
    pp_ms = st_pp["pp"] / max(n_pp, 1)
    pm = pm_and_sort_roofline(E, eng, box, {}, reps=2)["pm"]
    b0 = box
    d, k = E.init_drift_table(b0.a_start, b0.a_end, ics.OMEGA_M, ics.OMEGA_L,
                              ics.HUBBLE_INTERNAL)
    f = E.kdk_factors(d, k, b0.a_start, b0.a_end, b0.a_start, b0.a_end, steps)
    ts = []
    for rep in range(2):                  # the first run captures the graph
        eng.upload(pos, vel, mass)
        eng.compute_accelerations(E.FP32, True)
        eng.synchronize()
        t = time.perf_counter()
        eng.kdk_steps(f, E.FP32, use_graph=True)
        ts.append(time.perf_counter() - t)
    eng.close()
    del vel
    step_ms = ts[-1] / steps * 1e3
    log(f"[bench] halo512: {step_ms:.1f} ms per KDK step, pair kernel {pp_ms:.1f} ms")
    achieved = FLOPS_PER_INTERACTION * pairs_in / (pp_ms * 1e-3) / 1e12
    out = {
        "workload": (f"halo512: {box.n_side}^3 particles, {box.M}^3 FP32 mesh, "
                     f"{box.L / 1000:g} Mpc/h, 50% in {box.n_halos} NFW halos + 50% "
                     f"uniform; {steps} FP32 KDK steps from a = {box.a_start}, one "
                     "hipGraph per step"),
        "n_particles": box.N,
        "steps": steps,
        "ms_per_step": step_ms,
        "evals_per_s": box.N / (step_ms * 1e-3),
        "tuning": tuning,
        "pp": {"ms": pp_ms, "pairs_in": pairs_in, "candidates": candidates,
               "interactions_per_s": pairs_in / (pp_ms * 1e-3),
               "achieved_tflops": achieved, "peak_tflops": PEAK_FP32_TFLOPS,
               "frac": achieved / PEAK_FP32_TFLOPS},
        "pm_roofline": pm,
        "error_vs_oracle": None,
    }
    if with_oracle:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
        threads = min(threads, os.cpu_count() or threads)
        p = oracle.params(box.M, box.L, ics.G_INTERNAL, box.asmth, box.rcut,
                          box.soft_h, nthreads=threads)
        rng = np.random.default_rng(seed)
        targets = np.sort(rng.choice(box.N, n_oracle, replace=False))
        t0 = time.perf_counter()
        ref_pp, _ = oracle.pp_force(p, pos, mass, targets)
        t_pp = time.perf_counter() - t0
        log(f"[bench] halo512 oracle PP on {n_oracle} targets in {t_pp:.1f}s; PM next")
        t0 = time.perf_counter()
        ref_pm = oracle.pm_force(p, pos, mass)[targets]
        t_pm = time.perf_counter() - t0
        tot = ref_pp + ref_pm
        den = np.linalg.norm(tot, axis=1)
        den = np.where(den > 0, den, 1.0)
        a, q = a32[targets], p32[targets]
        out["error_vs_oracle"] = {
            "targets": n_oracle,
            "seed": seed,
            "denominator": "|a_total| of the oracle",
            "pp": percentiles(np.linalg.norm(a - ref_pp, axis=1) / den),
            "pm": percentiles(np.linalg.norm(q - ref_pm, axis=1) / den),
            "total": percentiles(np.linalg.norm(a + q - tot, axis=1) / den),
            "oracle_s": {"pp": t_pp, "pm": t_pm, "threads": threads},
        }
        log(f"[bench] halo512 error vs oracle {out['error_vs_oracle']['total']} "
            f"(oracle PP {t_pp:.1f}s, PM {t_pm:.1f}s)")
    return out


def pmc_traffic(kernel: str, workload: str, key: str = "traffic_bytes"):
    """HBM bytes per dispatch of `kernel` on `workload` from the committed
    rocprofv3 PMC summary (profiles/<kernel>_<workload>_pmc.md, the line
    `<key>: <B>`), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="halo256")
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-targets", type=int, default=100000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-boxes", type=int, default=64,
                    help="boxes of the batched-vs-looped ensemble run (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--timeline-steps", type=int, default=12,
                    help="individual-timestep steps timed after the first (0: skip)")
    ap.add_argument("--halo512-steps", type=int, default=5,
                    help="FP32 KDK steps of the 512^3 / 1024^3 leg (0: skip)")
    ap.add_argument("--halo512-oracle", type=int, default=1000000,
                    help="random particles of the 512^3 leg checked against the "
                         "FP64 oracle (0: no oracle)")
    ap.add_argument("--pm-overlap", choices=["auto", "0", "1"], default="auto",
                    help="PM chain on a second stream beside the pair force: "
                         "the autotuner's choice, or forced off / on")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="counter pass: one FP32 evaluation of the workload and "
                         "a 1 GiB device copy of known size, nothing else")
    ap.add_argument("--tuning", default=None,
                    help="skip the autotuner and use SHIFT,WG,OVERLAP,SYMMETRIC"
                         "[,CELL_LEVEL,FFT_SPLIT,WIDE_TILES] (e.g. 3,8,0,1): a run for "
                         "a clean kernel trace")
    ap.add_argument("--pp-symmetric", choices=["auto", "0", "1"], default="auto",
                    help="FP32 pair force evaluating each pair once: the "
                         "autotuner's choice, or forced off / on")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the ensemble leg and print its JSON object "
                         "(a run for a kernel trace of the batched ensemble)")
    ap.add_argument("--ensemble-looped", type=int, choices=[0, 1], default=1,
                    help="time the ensemble's boxes one by one as well (0: batched only)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not "
              "supported (the engine runs one GPU per handle; the whole job "
              "here is one GPU's share)", file=sys.stderr)
        return 2

    from cuco_amd import engine as E
    from cuco_amd import ics

    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_bench(E, ics, args.ensemble_boxes,
                                                     args.ensemble_steps,
                                                     looped=bool(args.ensemble_looped))}),
              flush=True)
        return 0

    box = ics.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    pos, vel, mass = ics.generate(box)
    log(f"[bench] {box.name}: N={box.N} M={box.M} generated in {time.perf_counter() - t0:.1f}s")

    eng = E.Engine(max_particles=box.N, **box.engine_kwargs())
    eng.upload(pos, vel, mass)
    log("[bench] uploaded")
    if args.pmc_probe:
        if args.pp_symmetric != "auto":
            t = eng.tuning()
            t["pp_symmetric"] = int(args.pp_symmetric)
            eng.set_tuning(**t)
        eng.compute_accelerations(E.FP32, True)
        eng.synchronize()
        nbytes = 1 << 30
        src, dst = E.DeviceBuffer(nbytes), E.DeviceBuffer(nbytes)
        dst.copy_from(src, nbytes)
        log("[bench] pmc probe done: one FP32 evaluation, one 1 GiB copy")
        eng.close()
        return 0

    # FP64 plain path (reference for the error percentiles).
    eng.compute_accelerations(E.FP64, True)
    eng.synchronize()
    ref_pp, ref_pm = eng.accelerations()
    eng.set_timing(True)
    eng.stage_times(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        eng.compute_accelerations(E.FP64, True)
    eng.synchronize()
    t64 = (time.perf_counter() - t0) / args.fp64_steps
    st64, _ = eng.stage_times(reset=True)
    log(f"[bench] fp64 {t64 * 1e3:.1f} ms/eval")

    # FP32 chosen path: autotune (outside the timed region), warmup, steps.
    t0 = time.perf_counter()
    if args.tuning:
        v = [int(x) for x in args.tuning.split(",")] + [0, 0, 0]
        tuning = dict(pp_block_shift=v[0], pp_workgroups_per_cu=v[1],
                      pm_overlap=v[2], pp_symmetric=v[3], pp_cell_level=v[4],
                      fft_split=v[5], pp_wide_tiles=v[6])
        eng.set_tuning(**tuning)
        log(f"[bench] tuning {tuning} (given)")
    else:
        tuning, tune_ms = eng.autotune()
        log(f"[bench] autotune {tuning} best pp {tune_ms:.2f} ms "
            f"({time.perf_counter() - t0:.1f}s)")
    other_ms = None
    if args.tuning:
        args.pm_overlap = str(tuning["pm_overlap"])
    if args.pp_symmetric != "auto":
        tuning["pp_symmetric"] = int(args.pp_symmetric)
        eng.set_tuning(**tuning)
    if args.pm_overlap != "auto":
        tuning["pm_overlap"] = int(args.pm_overlap)
        eng.set_tuning(**tuning)
    else:
        # The other PM/PP overlap setting, timed for comparison (not the value).
        other = dict(tuning, pm_overlap=1 - tuning["pm_overlap"])
        eng.set_tuning(**other)
        for _ in range(2):
            eng.compute_accelerations(E.FP32, True)
        eng.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(2, args.steps // 2)):
            eng.compute_accelerations(E.FP32, True)
        eng.synchronize()
        other_ms = (time.perf_counter() - t0) / max(2, args.steps // 2) * 1e3
        eng.set_tuning(**tuning)
        log(f"[bench] fp32 with pm_overlap={other['pm_overlap']}: {other_ms:.2f} ms/eval")
    for _ in range(args.warmup):
        eng.compute_accelerations(E.FP32, True)
    eng.synchronize()
    eng.stage_times(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.compute_accelerations(E.FP32, True)
    eng.synchronize()
    t32 = time.perf_counter() - t0
    st32, nev = eng.stage_times(reset=True)
    eng.set_timing(False)
    ms32 = t32 / args.steps * 1e3
    log(f"[bench] fp32 {ms32:.2f} ms/eval over {args.steps} steps")

    a_pp, a_pm = eng.accelerations()
    tot = ref_pp + ref_pm
    den = np.linalg.norm(tot, axis=1)
    den = np.where(den > 0, den, 1.0)
    errors = {
        "pp": percentiles(np.linalg.norm(a_pp - ref_pp, axis=1) / den),
        "pm": percentiles(np.linalg.norm(a_pm - ref_pm, axis=1) / den),
        "total": percentiles(np.linalg.norm(a_pp + a_pm - tot, axis=1) / den),
    }
    log(f"[bench] errors {errors['total']}")

    pairs_in, candidates = eng.count_pairs()
    # The pair kernel alone (no PM running beside it), HIP events on the
    # engine's stream.
    eng.set_timing(True)
    eng.stage_times(reset=True)
    for _ in range(3):
        eng.gravity_linklist(E.FP32)
    eng.synchronize()
    st_pp, n_pp = eng.stage_times(reset=True)
    eng.set_timing(False)
    pp_ms = st_pp["pp"] / max(n_pp, 1)
    achieved = FLOPS_PER_INTERACTION * pairs_in / (pp_ms * 1e-3) / 1e12
    roofline = {
        "bound": "valu_fp32",
        "note": ("compute roofline of the FP32 pair kernel: FP32 vector (VALU) "
                 "instructions, no MFMA (a pair force is no dense contraction), "
                 "against MI355X's published FP32 vector peak; packed FP32 is "
                 "disabled in every kernel, which caps this kernel at half of "
                 "that peak"),
        "kernel": ("k_pp32 (FP32 pair force, each pair once for both particles) + "
                   "k_pp32_combine" if tuning.get("pp_symmetric")
                   else "k_pp32 (FP32 pair force)"),
        "achieved": achieved,
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32_TFLOPS,
        "traffic": pmc_traffic("pp32", box.name,
                               "traffic_bytes_symmetric" if tuning.get("pp_symmetric")
                               else "traffic_bytes"),
        "traffic_note": (f"HBM bytes per launch from the committed rocprofv3 "
                         f"FETCH_SIZE / WRITE_SIZE passes (profiles/pp32_{box.name}_pmc.md: "
                         f"8^3-cell blocks, 8 workgroups per CU, tiles of 64 targets), "
                         f"not from this run"),
        "pp_ms": pp_ms,
        "pairs_in": pairs_in,
        "candidates": candidates,
        "flops_per_interaction": FLOPS_PER_INTERACTION,
    }

    value = box.N * args.steps / t32
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle force evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{box.name}: {box.n_side}^3 particles, {box.M}^3 mesh, "
                         f"{box.L / 1000:g} Mpc/h, 50% in {box.n_halos} NFW halos "
                         "+ 50% uniform; one complete PP+PM force evaluation per step"),
            "n_particles": box.N,
            "pm_grid": box.M,
            "path": "fp32 cell-relative pairs + fp32 mesh, fp64 positions",
        },
        "fp64_value": box.N / t64,
        "fp64_ms_per_step": t64 * 1e3,
        "fp32_over_fp64": t64 * 1e3 / ms32,
        "error_vs_fp64": errors,
        "tuning": tuning,
        "fp32_ms_other_overlap": other_ms,
        "stage_ms_fp32": {k: v / max(nev, 1) for k, v in st32.items()},
        "stage_ms_note": ("HIP-event time per stage; with pm_overlap=1 the PM stages "
                          "(deposit .. interp) run on a second stream beside the pair "
                          "force, so their times include waiting for the compute units "
                          "the pair kernel holds" if tuning.get("pm_overlap") else
                          "HIP-event time per stage, one stream"),
        "stage_ms_fp64": {k: v / max(args.fp64_steps, 1) for k, v in st64.items()},
        "roofline": roofline,
        "pm_roofline": pm_and_sort_roofline(E, eng, box, {k: v / max(nev, 1)
                                                          for k, v in st32.items()}),
        "cpu_baseline": None,
    }
    eng.close()
    if args.timeline_steps > 0:
        out["timeline"] = timeline_bench(E, ics, box, pos, vel, mass, args.timeline_steps)
    if args.ensemble_boxes > 0:
        out["ensemble"] = ensemble_bench(E, ics, args.ensemble_boxes,
                                         args.ensemble_steps,
                                         looped=bool(args.ensemble_looped))
    if args.halo512_steps > 0 and box.name != "halo512":
        out["halo512"] = halo512_bench(E, ics, tuning, args.halo512_steps,
                                       args.halo512_oracle,
                                       with_oracle=args.halo512_oracle > 0)
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        parity, out["cpu_baseline"] = cpu_baseline(box, pos, mass, args.cpu_targets,
                                                   gpu_fp64=(ref_pp, ref_pm))
        out["fp64_vs_oracle"] = parity
        log(f"[bench] cpu baseline in {time.perf_counter() - t0:.1f}s; FP64 vs oracle {parity}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
