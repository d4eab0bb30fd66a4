"""Throughput benchmark of the particle-filter hot path on one MI355X.

Workload (one GPU): a single synthetic sequence, K=15 components, 12-dim PCA
state, 6-dim measurement (head + two hands), N = 100,000 particles, 8
candidate detections per hand (1 true + Poisson(7) clutter, PDAF with a
chi-square 99% gate), FP32 state, FP64 log-likelihood accumulation, Philox
uniforms shared with the CPU oracle.  One step = one frame of the filter over
all N particles; the timed region is device-resident (inputs in HBM).

Prints one JSON line with the metric, the roofline of the dominant kernel
(the fused step kernel, HBM-bound) and the CPU baseline (the FP64 oracle on a
bounded sample of the same workload, on this host).

``--workload batched`` instead runs many ragged sequences (lengths
U{100..2000} frames, N log-uniform in [1k, 64k], own GMM each) through the
batched engine and reports the default and the autotuned schedule.
``--workload adjoint`` times the gradient of the log-marginal-likelihood of
64 sequences x 500 frames x 16k particles (checkpoint every 50 frames)
against the forward pass.
``--workload scaled`` runs the 24-dimensional state: K=64, m=10 (5 joints,
PDAF on each, 16 candidates, Poisson(10) clutter), 32 sequences x 31,250
particles x 1000 frames through the batched engine, with the covariance tiles
on the matrix cores (the line's value) beside the one-particle-per-lane and
four-lane vector-ALU kernels.
``--workload tracker`` runs the pose tracker's whole frame for both arms
(likelihood blur, 10 N hand proposals per arm, association weights, proposal
resampling, measurements, both filter updates, pose output) with the shipped
GMM models at N = 100,000 particles per arm.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def step_bytes_per_particle(d: int) -> int:
    """Algorithmic HBM bytes of one fused step per particle: read the
    ancestor's x and packed P (FP32), write the new x and P, write the
    log-weight, read the ancestor index."""
    np_ = d * (d + 1) // 2
    return 8 * (d + np_) + 4 + 4


def frame_bytes_per_particle(d: int) -> int:
    """Algorithmic HBM bytes per particle of the single-sequence step that
    also writes its frame's resampler weights: the fused step, plus the
    block-local scan of the fixed-point weight written (8) and read back by
    the next frame's ancestor search (8)."""
    return step_bytes_per_particle(d) + 16


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--particles", type=int, default=100_000)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="default workload: skip the adjoint measurement")
    ap.add_argument("--block", type=int, default=0, help="step kernel block size (0 = default)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run a device-to-device copy of this size (PMC calibration)")
    ap.add_argument("--workload", choices=("single", "batched", "adjoint", "scaled", "tracker"), default="single",
                    help="single: one sequence of N particles (the headline line); batched: many ragged "
                         "sequences in one persistent step per frame, tuned vs default schedule")
    ap.add_argument("--sequences", type=int, default=256, help="batched: number of sequences")
    ap.add_argument("--tune-frames", type=int, default=32, help="batched: frames timed per candidate schedule")
    ap.add_argument("--scaled-variants", type=str, default="valu,quad_valu,mfma",
                    help="scaled: kernels to run, of valu, quad_valu, mfma (the line's value is mfma's)")
    ap.add_argument("--scaled-shape", type=str, default="32,31250,1000",
                    help="scaled: sequences,particles,frames (the measured workload is 32,31250,1000)")
    ap.add_argument("--tracker-shape", type=str, default="100000,400",
                    help="tracker: particles per arm,frames (the measured workload is 100000,400)")
    ap.add_argument("--adj-shape", type=str, default="64,16384,500",
                    help="adjoint: sequences,particles,frames (the measured workload is 64,16384,500)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "adjoint":
        return run_adjoint(args)
    if args.workload == "scaled":
        return run_scaled(args)
    if args.workload == "tracker":
        return run_tracker(args)

    from mkfbodytracker_pdaf_amd import synthetic
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray, synchronize
    from mkfbodytracker_pdaf_amd.rbpf import Rbpf

    N, K, W = args.particles, args.steps, args.warmup
    T = K + W
    cap, cpj = 16, 8
    lam = (cpj - 1) / (640.0 * 480.0)
    seed = 2024
    mdl = synthetic.make_model(seed=seed)
    seq = synthetic.simulate(mdl, T, seed=seed, cand_per_joint=cpj, cap=cap)

    rb = Rbpf(N, cand_capacity=cap, clutter_density=lam, block_size=args.block)
    rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean,
                  sel=mdl.sel, R=mdl.R)
    rb.set_joint_clutter(mdl.clutter)
    bins = synthetic.init_bins(mdl, N, seed, lambda w, n, U: rb.resample(w, n, U)[0])
    rb.reset(bins)

    d_count = DeviceArray(seq.count.astype(np.int32))
    d_cand = DeviceArray(seq.cand.astype(np.float32))
    d_uni = DeviceArray(seq.uniforms.astype(np.uint32))
    d_lml = DeviceArray.empty(T, np.float64)
    flags = 0 if args.no_graph else 1

    # warmup frames (also JIT-free: the library is prebuilt)
    uni_off = d_uni.ptr
    rb.run_candidates_dev(W, d_count.ptr, d_cand.ptr, uni_off, d_lml.ptr, flags)
    rb.synchronize()
    # timed region: frames W .. W+K-1, inputs already resident in HBM
    stride_cnt = seq.count.shape[1] * 4
    stride_cand = seq.cand.shape[1] * cap * 2 * 4
    c_ptr = d_count.ptr + W * stride_cnt
    k_ptr = d_cand.ptr + W * stride_cand
    u_ptr = d_uni.ptr + W * 2 * 4
    l_ptr = d_lml.ptr + W * 8
    synchronize()
    t0 = time.perf_counter()
    rb.run_candidates_dev(K, c_ptr, k_ptr, u_ptr, l_ptr, flags)
    rb.synchronize()
    synchronize()
    dt = time.perf_counter() - t0
    lml = d_lml.get()[W:]

    graph_frames = rb.last_graph_frames()

    # kernel timing over the same frames, HIP events on the handle's stream:
    # one pair around the K frames of the run (one k_step launch per frame,
    # replayed from graphs unless --no-graph), then a pair around every launch
    # without graphs (each launch's own dispatch included), reported beside it
    rb.reset(bins)
    rb.run_candidates_dev(K, c_ptr, k_ptr, u_ptr, l_ptr, flags | 4)
    rb.synchronize()
    step_ms_total, _, launches = rb.last_kernel_time()
    timed_graph_frames = rb.last_graph_frames()
    step_avg_s = step_ms_total / launches / 1e3
    rb.reset(bins)
    rb.run_candidates_dev(K, c_ptr, k_ptr, u_ptr, l_ptr, 2)
    rb.synchronize()
    step_ms_ng, frame_ms_total, launches_ng = rb.last_kernel_time()
    bpp = frame_bytes_per_particle(mdl.d)
    achieved = bpp * N / step_avg_s / 1e9

    calib_copy(args.calib_copy_mb)

    traffic = measured_traffic(N, mdl.d, mdl.K)
    evals = N * K
    out = {
        "metric": "likelihood evals/s at N=100k particles (single sequence, PDAF candidates)",
        "value": evals / dt,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": dt / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded GMM K=15 d=12, pinhole-projected 3-joint pose, Philox uniforms)",
        "config": {"workload": "single-sequence throughput: K=15, d=12, m=6, N=%d, %d frames, "
                               "8 candidates per hand (1 true + Poisson(7) clutter), chi2 99%% gate" % (N, K),
                   "particles": N, "frames": K, "state_dim": mdl.d, "components": mdl.K,
                   "graph": graph_frames == K, "graph_frames": graph_frames},
        "frames_per_s": K / dt,
        "lml_total": float(np.sum(lml)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic["bytes_per_launch"],
                     "traffic_source": None if traffic is None else
                     "profiles/k_step_cand_n100k.md (rocprofv3 FETCH_SIZE x %.2f + WRITE_SIZE x %.2f, "
                     "copy-kernel calibration)" % (traffic["calib_fetch"], traffic["calib_write"]),
                     "algorithmic_bytes": bpp * N,
                     "kernel": "k_step<12,cand>", "bytes_per_particle": bpp,
                     "avg_launch_us": step_avg_s * 1e6, "launches": launches,
                     "timing": "one HIP event pair on the handle's stream around the %d frames of a run "
                               "(%d of them replayed from graphs): frame time, launch gaps included" %
                               (launches, timed_graph_frames),
                     "avg_launch_us_no_graph": step_ms_ng / launches_ng * 1e3,
                     "frame_us_no_graph": frame_ms_total / launches_ng * 1e3},
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mdl, seq, N, cap, lam, args.cpu_seconds)
        out["cpu_baseline"]["all_cores"] = cpu_baseline(mdl, seq, N, cap, lam, args.cpu_seconds,
                                                        threads=host_cores())
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
        out["speedup_vs_cpu_all_cores"] = out["value"] / out["cpu_baseline"]["all_cores"]["value"]
    if not args.no_adjoint:
        # the adjoint workload, measured after the headline's timed region
        rb.close()
        m = measure_adjoint(args.adj_shape)
        out["adjoint_over_forward"] = m["t_grad"] / m["t_fwd"]
        out["adjoint"] = {"workload": "gradient of the summed log-marginal-likelihood: %d sequences x %d frames x "
                                      "N=%d, K=15, d=12, checkpoint every %d frames" % (m["S"], m["T"], m["N"], m["C"]),
                          "forward_s": m["t_fwd"], "gradient_s": m["t_grad"],
                          "gradient_recorded_forward_ms": m["rec_fwd_ms"], "gradient_backward_ms": m["bwd_ms"],
                          "lml_identical_in_gradient_run": bool(np.array_equal(m["lml_fwd"], m["lml_grad"])),
                          "timing": "wall clock of the second of two runs each, after the headline's timed region"}
    print(json.dumps(out))


def calib_copy(mb: int) -> None:
    """One device-to-device copy of mb MiB (the byte counters' calibration
    under rocprofv3 --pmc); nothing when mb is 0."""
    if not mb:
        return
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray, synchronize
    nbytes = mb << 20
    a = DeviceArray(np.zeros(nbytes // 4, np.float32))
    b = DeviceArray.empty(nbytes // 4, np.float32)
    b.copy_from(a)
    synchronize()
    a.close()
    b.close()


def host_cores() -> int:
    """Cores this process may use, at most 16 (the CPU share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def measured_traffic(N, d, K):
    """HBM bytes per step-kernel launch measured with rocprofv3 counters on
    this workload (the committed summary next to the profile), or None for
    other sizes."""
    path = os.path.join(ROOT, "profiles", "k_step_cand_n100k_traffic.json")
    if (N, d, K) != (100_000, 12, 15) or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def run_batched(args):
    """Many independent sequences (lengths U{100..2000} frames, particle counts
    log-uniform in [1k, 64k], each with its own seeded GMM) advanced together,
    one persistent step kernel and one weights kernel per frame, the frame
    loop replayed from a hipGraph.  Times the whole run with the default
    schedule, autotunes, then times the whole run again with the tuned one."""
    from mkfbodytracker_pdaf_amd import synthetic
    from mkfbodytracker_pdaf_amd.batch import RbpfBatch
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray, synchronize
    from mkfbodytracker_pdaf_amd.rbpf import Rbpf

    seed, S, cap, cpj = 7, args.sequences, 16, 8
    lam = (cpj - 1) / (640.0 * 480.0)
    rng = np.random.default_rng(seed)
    Ts = rng.integers(100, 2001, S).astype(np.int32)
    Ns = np.exp(rng.uniform(np.log(1000), np.log(64000), S)).astype(np.int32)
    models = [synthetic.make_model(seed=seed * 1000 + s) for s in range(S)]
    seqs = [synthetic.simulate_fast(models[s], int(Ts[s]), seed=seed * 1000 + s, cand_per_joint=cpj, cap=cap)
            for s in range(S)]
    b = RbpfBatch(Ns, Ts, K=models[0].K, m=models[0].m, cand_capacity=cap, clutter_density=lam)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    helper = Rbpf(int(Ns.max()))
    bins = [synthetic.init_bins(models[s], int(Ns[s]), seed * 1000 + s,
                                lambda w, n, U: helper.resample(w, n, U)[0]) for s in range(S)]
    helper.close()

    def reset_all():
        for s in range(S):
            b.reset(s, bins[s])

    d_cnt = DeviceArray(np.concatenate([q.count for q in seqs]).astype(np.int32))
    d_cand = DeviceArray(np.concatenate([q.cand for q in seqs]).astype(np.float32))
    d_uni = DeviceArray(np.concatenate([q.uniforms for q in seqs]).astype(np.uint32))
    F = int(Ts.sum())
    d_lml = DeviceArray.empty(F, np.float64)

    def timed_run():
        reset_all()
        synchronize()
        t0 = time.perf_counter()
        b.run_dev(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, graph=not args.no_graph)
        synchronize()
        return time.perf_counter() - t0, b.last_run_ms() / 1e3

    default = (256, 1, 2)
    b.schedule = default
    timed_run()                                   # warm-up run
    t_def, dev_def = timed_run()
    lml_def = d_lml.get()
    reset_all()
    best, tune_best_ms, tune_def_ms = b.autotune(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr,
                                                 frames=args.tune_frames)
    t_tuned, dev_tuned = timed_run()
    lml_tuned = d_lml.get()

    evals = float(np.sum(Ns.astype(np.int64) * Ts))
    bpp = step_bytes_per_particle(models[0].d)
    achieved = bpp * evals / t_tuned / 1e9
    out = {
        "metric": "likelihood evals/s, batched ragged sequences (PDAF candidates)",
        "value": evals / t_tuned,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": int(Ts.max()),
        "warmup": 1,
        "ms_per_step": t_tuned / int(Ts.max()) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (per-sequence seeded GMM K=15 d=12, pinhole-projected 3-joint pose, Philox uniforms)",
        "config": {"workload": "batched: %d sequences, lengths U{100..2000} frames, N log-uniform 1k..64k, "
                               "K=15, d=12, m=6, 8 candidates per hand" % S,
                   "sequences": S, "particle_frames": evals, "sequence_frames": F,
                   "particles_total": int(Ns.sum()), "frames_max": int(Ts.max())},
        "sequence_frames_per_s": F / t_tuned,
        "schedule_default": {"block": default[0], "ppl": default[1], "blocks_per_cu": default[2],
                             "wall_s": t_def, "device_s": dev_def, "evals_per_s": evals / t_def},
        "schedule_tuned": {"block": best[0], "ppl": best[1], "blocks_per_cu": best[2],
                           "wall_s": t_tuned, "device_s": dev_tuned, "evals_per_s": evals / t_tuned,
                           "tuning_frames": args.tune_frames, "tuning_ms_best": tune_best_ms,
                           "tuning_ms_default": tune_def_ms},
        "tuned_over_default": t_def / t_tuned,
        "lml_total": float(np.sum(lml_tuned)),
        "lml_identical_across_schedules": bool(np.array_equal(lml_def, lml_tuned)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_batch_step + k_batch_weights (whole run)", "bytes_per_particle": bpp},
    }
    if not args.no_cpu:
        s0 = int(np.argmax(Ns))
        out["cpu_baseline"] = cpu_baseline(models[s0], seqs[s0], int(Ns[s0]), cap, lam, args.cpu_seconds,
                                           seed=seed * 1000 + s0)
        out["speedup_vs_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out))


def measure_adjoint(shape: str) -> dict:
    """Gradient of the summed log-marginal-likelihood of S sequences x T
    frames x N particles (K=15, d=12, own GMM per sequence) w.r.t. means,
    logit weights, reversion rates and Cholesky factors, checkpointing every
    50 frames, and the plain (hipGraph) forward pass of the same inputs; the
    second of two runs of each is timed."""
    from mkfbodytracker_pdaf_amd import synthetic
    from mkfbodytracker_pdaf_amd.batch import RbpfBatch
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray, synchronize
    from mkfbodytracker_pdaf_amd.rbpf import Rbpf

    S, N, T = (int(v) for v in shape.split(","))
    seed, C, cap, cpj = 11, 50, 16, 8
    lam = (cpj - 1) / (640.0 * 480.0)
    models = [synthetic.make_model(seed=seed * 1000 + s) for s in range(S)]
    seqs = [synthetic.simulate_fast(models[s], T, seed=seed * 1000 + s, cand_per_joint=cpj, cap=cap)
            for s in range(S)]
    b = RbpfBatch([N] * S, [T] * S, K=models[0].K, m=models[0].m, cand_capacity=cap, clutter_density=lam)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    helper = Rbpf(N)
    bins = [synthetic.init_bins(models[s], N, seed * 1000 + s, lambda w, n, U: helper.resample(w, n, U)[0])
            for s in range(S)]
    helper.close()

    def reset_all():
        for s in range(S):
            b.reset(s, bins[s])

    d_cnt = DeviceArray(np.concatenate([q.count for q in seqs]).astype(np.int32))
    d_cand = DeviceArray(np.concatenate([q.cand for q in seqs]).astype(np.float32))
    d_uni = DeviceArray(np.concatenate([q.uniforms for q in seqs]).astype(np.uint32))
    d_lml = DeviceArray.empty(S * T, np.float64)
    # plain forward (warm-up, then timed)
    for rep in range(2):
        reset_all()
        synchronize()
        t0 = time.perf_counter()
        b.run_dev(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, graph=True)
        synchronize()
        t_fwd = time.perf_counter() - t0
    lml_fwd = d_lml.get()
    # gradient (warm-up allocates the records, then timed)
    for rep in range(2):
        reset_all()
        synchronize()
        t0 = time.perf_counter()
        g = b.lml_grad(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, checkpoint_every=C)
        synchronize()
        t_grad = time.perf_counter() - t0
    rec_fwd_ms, bwd_ms = b.last_adjoint_ms()
    lml_grad = d_lml.get()
    b.close()
    return dict(S=S, N=N, T=T, C=C, t_fwd=t_fwd, t_grad=t_grad, rec_fwd_ms=rec_fwd_ms, bwd_ms=bwd_ms,
                lml_fwd=lml_fwd, lml_grad=lml_grad, g=g)


def run_adjoint(args):
    """The adjoint workload (measure_adjoint) as its own bench line."""
    calib_copy(args.calib_copy_mb)
    m = measure_adjoint(args.adj_shape)
    S, N, T, C = m["S"], m["N"], m["T"], m["C"]
    t_fwd, t_grad, rec_fwd_ms, bwd_ms = m["t_fwd"], m["t_grad"], m["rec_fwd_ms"], m["bwd_ms"]
    lml_fwd, lml_grad, g = m["lml_fwd"], m["lml_grad"], m["g"]
    evals = float(S) * N * T
    out = {
        "metric": "adjoint gradient time / forward time (log-marginal-likelihood, batched)",
        "value": t_grad / t_fwd,
        "unit": "x forward",
        "n_gpus": 1,
        "steps": T,
        "warmup": 1,
        "ms_per_step": t_grad / T * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (per-sequence seeded GMM K=15 d=12, pinhole-projected 3-joint pose, Philox uniforms)",
        "config": {"workload": "adjoint: %d sequences x %d frames x N=%d, K=15, d=12, m=6, 8 candidates per hand, "
                               "checkpoint every %d frames" % (S, T, N, C),
                   "sequences": S, "frames": T, "particles": N, "checkpoint_every": C},
        "forward_s": t_fwd,
        "forward_evals_per_s": evals / t_fwd,
        "gradient_s": t_grad,
        "gradient_recorded_forward_ms": rec_fwd_ms,
        "gradient_backward_ms": bwd_ms,
        "backward_over_recorded_forward": bwd_ms / rec_fwd_ms,
        "lml_total": float(np.sum(lml_fwd)),
        "lml_identical_in_gradient_run": bool(np.array_equal(lml_fwd, lml_grad)),
        "grad_norms": {k: float(np.linalg.norm(v)) for k, v in g.items()},
    }
    print(json.dumps(out))


def run_scaled(args):
    """24-dimensional PCA state, K=64 components, 10-dimensional measurement (5
    joints in 2-D, every joint PDAF with 1 true + Poisson(10) clutter
    detections, capacity 16), 32 sequences x 31,250 particles x 1000 frames,
    one step kernel and one weights kernel per frame for all sequences."""
    from mkfbodytracker_pdaf_amd import synthetic
    from mkfbodytracker_pdaf_amd.batch import RbpfBatch
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray, synchronize
    from mkfbodytracker_pdaf_amd.rbpf import Rbpf

    S, N, T = (int(v) for v in args.scaled_shape.split(","))
    seed, cap, cpj, K, d = 13, 16, 11, 64, 24
    lam = (cpj - 1) / (640.0 * 480.0)
    joints = ("head", "left_hand", "right_hand", "left_elbow", "right_elbow")
    models = [synthetic.make_model(seed=seed * 1000 + s, K=K, d=d, joints=joints, clutter=(1, 1, 1, 1, 1))
              for s in range(S)]
    seqs = [synthetic.simulate_fast(models[s], T, seed=seed * 1000 + s, cand_per_joint=cpj, cap=cap) for s in range(S)]
    b = RbpfBatch([N] * S, [T] * S, K=K, m=10, d=d, cand_capacity=cap, clutter_density=lam)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    helper = Rbpf(N)
    bins = [synthetic.init_bins(models[s], N, seed * 1000 + s, lambda w, n, U: helper.resample(w, n, U)[0])
            for s in range(S)]
    helper.close()
    d_cnt = DeviceArray(np.concatenate([q.count for q in seqs]).astype(np.int32))
    d_cand = DeviceArray(np.concatenate([q.cand for q in seqs]).astype(np.float32))
    d_uni = DeviceArray(np.concatenate([q.uniforms for q in seqs]).astype(np.uint32))
    d_lml = DeviceArray.empty(S * T, np.float64)
    from mkfbodytracker_pdaf_amd import capi
    names = {capi.PF_BATCH_MFMA: "mfma", capi.PF_BATCH_QUAD_VALU: "quad_valu", capi.PF_BATCH_VALU: "valu"}
    evals = float(S) * N * T
    bpp = step_bytes_per_particle(d)
    variants = {}
    by_name = {n: v for v, n in names.items()}
    chosen = [by_name[n] for n in args.scaled_variants.split(",")]
    for v in chosen:
        b.variant = v
        for rep in range(2):                      # warm-up run, then the timed run
            for s in range(S):
                b.reset(s, bins[s])
            synchronize()
            t0 = time.perf_counter()
            b.run_dev(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, graph=not args.no_graph)
            synchronize()
            dt = time.perf_counter() - t0
        variants[names[v]] = {"evals_per_s": evals / dt, "wall_s": dt, "device_s": b.last_run_ms() / 1e3,
                              "ms_per_frame": dt / T * 1e3, "hbm_frac": bpp * evals / dt / 1e9 / HBM_PEAK_GBS,
                              "lml_total": float(np.sum(d_lml.get()))}
    calib_copy(args.calib_copy_mb)
    head = "mfma" if "mfma" in variants else next(iter(variants))
    dt = variants[head]["wall_s"]
    achieved = bpp * evals / dt / 1e9
    out = {
        "metric": "likelihood evals/s, scaled state (d=24, K=64, 5 PDAF joints), batched",
        "value": evals / dt,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": T,
        "warmup": 1,
        "ms_per_step": dt / T * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (per-sequence seeded GMM K=64 d=24, pinhole-projected 5-joint pose, Philox uniforms)",
        "config": {"workload": "scaled state: %d sequences x %d particles x %d frames, K=64, d=24, m=10, "
                               "16 candidates per joint (1 true + Poisson(10) clutter), FP32 matrix cores "
                               "(v_mfma_f32_4x4x1f32, four lanes per particle)" % (S, N, T),
                   "sequences": S, "particles": N, "frames": T, "state_dim": d, "components": K},
        "frames_per_s": T / dt,
        "lml_total": variants[head]["lml_total"],
        "variants": variants,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_batch_step_q24<5, %s> + k_batch_weights (whole run)" % head,
                     "bytes_per_particle": bpp},
    }
    if "mfma" in variants:
        for other in ("valu", "quad_valu"):
            if other in variants:
                out["mfma_over_" + other] = variants[other]["wall_s"] / variants["mfma"]["wall_s"]
    print(json.dumps(out))


def tracker_bytes_per_particle(d: int = 12, ppp: int = 10) -> int:
    """Algorithmic HBM bytes per particle and arm of one tracker frame:
    ppp proposals (FP64 weight and block-local scan written, the scan read
    back by the selection), the measurement build (chosen index and 6-row
    column written), the step (plus its measurement column read), the frame
    weights' scan written by the step and read back by the ancestor search,
    the ancestor written, the resampled state read by the estimate."""
    return ppp * 24 + (4 + 24) + (frame_bytes_per_particle(d) + 24) + 4 + 4 * d


def run_tracker(args):
    """PFTracker's frame (pfPose.cpp:210-349) for both arms on the device: the
    shipped GMM files (left and right arm), N particles per arm, 10 N hand
    proposals per arm, synthetic 480 x 640 likelihood images with moving
    hand blobs and clutter, face ROIs; T frames in one device-resident run
    (16 captured frames replayed at a time)."""
    from mkfbodytracker_pdaf_amd import synthetic
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray, synchronize
    from mkfbodytracker_pdaf_amd.tracker import PoseTracker

    N, T = (int(v) for v in args.tracker_shape.split(","))
    left = os.path.join(ROOT, "tests", "golden", "data13D_PCA_100000_15_12.yml")
    right = os.path.join(ROOT, "tests", "golden", "data23D_PCA_100000_15_12.yml")
    distinct = min(T, 64)
    imgs, rois = synthetic.tracker_sequence(distinct, seed=21)
    reps = (T + distinct - 1) // distinct
    d_img = DeviceArray(np.tile(imgs, (reps, 1, 1))[:T])
    d_roi = DeviceArray(np.tile(rois, (reps, 1))[:T])
    d_lml = DeviceArray.empty(2 * T, np.float64)
    d_pose = DeviceArray.empty(T * 2 * 22, np.float64)
    d_p3 = DeviceArray.empty(T * 2 * 15, np.float64)
    for rep in range(2):                          # warm-up run, then the timed run
        tr = PoseTracker.from_yml(left, right, n_particles=N, seed=21)
        synchronize()
        t0 = time.perf_counter()
        tr.run_dev(T, d_img.ptr, d_roi.ptr, d_lml.ptr, d_pose.ptr, d_p3.ptr, graph=not args.no_graph)
        tr.synchronize()
        dt = time.perf_counter() - t0
        dev_ms = tr.last_run_ms()
        n_c = tr.n_c
        tr.close()
    calib_copy(args.calib_copy_mb)
    evals = 2.0 * N * T
    bpp = tracker_bytes_per_particle()
    achieved = bpp * evals / (dev_ms * 1e-3) / 1e9
    lml = d_lml.get().reshape(2, T)
    out = {
        "metric": "likelihood evals/s, pose tracker frame (both arms, proposals + association + updates + pose)",
        "value": evals / dt,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": T,
        "warmup": 1,
        "ms_per_step": dt / T * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 proposals / f32 state",
        "data": "shipped GMM models (data13D/data23D), synthetic likelihood images and face ROIs, Philox words",
        "config": {"workload": "pose tracker: 2 arms x N=%d particles, %d proposals per arm per frame, "
                               "480 x 640 likelihood, %d frames" % (N, n_c, T),
                   "particles_per_arm": N, "proposals_per_arm": n_c, "frames": T},
        "frames_per_s": T / dt,
        "proposals_per_s": 2.0 * n_c * T / dt,
        "device_ms_per_frame": dev_ms / T,
        "lml_mean": float(np.mean(lml)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole frame (all kernels of both arms, device events)", "bytes_per_particle": bpp},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_tracker(left, right, N, imgs, rois, args.cpu_seconds)
    print(json.dumps(out))


def cpu_baseline_tracker(left, right, N, imgs, rois, budget_s, seed=21):
    """The FP64 oracle of the tracker frame (1 thread): blur, proposals,
    weights, proposal resampling, measurements and both arm updates, N
    particles per arm, frames until the budget is spent (at least 1)."""
    import oracle
    from mkfbodytracker_pdaf_amd import capi
    oracle.set_threads(1)
    arms, Hp, M, states = [], [], [], []
    for a, path in enumerate((left, right)):
        means = capi.yml_read_matrix(path, "means")
        covs = capi.yml_read_matrix(path, "covs")
        w = capi.yml_read_matrix(path, "weights").ravel()
        g = capi.yml_read_matrix(path, "gamma").ravel()
        P = capi.yml_read_matrix(path, "pca_proj")
        m = capi.yml_read_matrix(path, "pca_mean").ravel()
        K, d = means.shape
        kf = oracle.load_gaussian(means[0], covs[:d], P, m, g[0], np.array([9, 10, 0, 1, 12, 13], np.int32),
                                  100.0 * np.eye(6))
        mdl = oracle.Model(w, g, means, covs.reshape(K, d, d), kf["H"], kf["BH"], kf["R"])
        bins, _ = oracle.resample_linear(w, N, int(oracle.philox_u32(seed, 32 + a, 0, 1)[0]))
        arms.append(mdl); Hp.append(P.T); M.append(m)
        states.append(mdl.reset(bins))
    orc = oracle.PoseTrackerOracle(arms, Hp, M, N, seed=seed)
    frames = 0
    t0 = time.perf_counter()
    while True:
        g = frames
        xbars = [oracle.rbpf_estimate(x) for x, _ in states]
        st = orc.proposal_stage(xbars, imgs[g % len(imgs)], rois[g % len(rois)], g, uniform=(g == 0))
        res = orc.update_arms(states, st["Z"], g)
        states = [(r["x"], r["P"]) for r in res]
        orc.pose([oracle.rbpf_estimate(x) for x, _ in states])
        frames += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return {"value": 2.0 * N * frames / el, "unit": "evals/s", "cores": 1, "kind": "port",
            "sample": f"{frames} tracker frames x 2 arms x N={N} (10 N proposals per arm), FP64 oracle, 1 thread",
            "cpu": _cpu_model(), "host_nproc": os.cpu_count()}


def cpu_baseline(mdl, seq, N, cap, lam, budget_s, seed=2024, threads=1):
    """FP64 oracle on the same workload: same model, the same frames'
    detections and uniforms, N particles, as many frames of the sequence as fit
    in the budget (at least 2); `threads` OpenMP threads over the particles."""
    import oracle
    oracle.set_threads(threads)
    om = oracle.Model(mdl.weights, mdl.gamma, mdl.means, mdl.Sigma, mdl.H(), mdl.BH(), mdl.R)
    from mkfbodytracker_pdaf_amd import synthetic
    bins = synthetic.init_bins(mdl, N, seed, lambda w, n, U: oracle.resample_linear(w, n, U)[0])
    x, P = om.reset(bins)
    frames = 0
    t0 = time.perf_counter()
    while True:
        t = frames
        U1, U2 = int(seq.uniforms[t, 0]), int(seq.uniforms[t, 1])
        r = om.update_pdaf(x, P, mdl.clutter, seq.count[t], seq.cand[t], U1, U2, 0.95,
                           9.210340371976184, lam)
        x, P = r["x"], r["P"]
        frames += 1
        el = time.perf_counter() - t0
        if frames >= 2 and (el >= budget_s or frames >= seq.count.shape[0]):
            break
    oracle.set_threads(1)
    return {"value": N * frames / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"{frames} frames x N={N} of the same sequence, FP64 oracle, {threads} thread(s)",
            "seconds": el, "cpu": _cpu_model(), "host_nproc": os.cpu_count()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
