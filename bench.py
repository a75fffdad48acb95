#!/usr/bin/env python3
"""Benchmark of the k-space SBE time loop on one MI355X.

Workload (default): the plain FP64 path at README scale -- 2-band
Haldane-like honeycomb (tests/golden/haldane_tb.dat, regenerated by
hhgamd.tbgen), 200x200 k-grid, nr3 = 1, complex128 rho, dt = 0.02197,
tmax = 90 (4096 RK4 steps), lambda = 3000 nm, I = 1e12 W/cm^2, sin^2 envelope
to 80, polarisation along y, CEP 0.

A "step" is one full solve over the resident inputs: 4096 RK4 steps of all
40 000 k-points plus the reduction of J(t) and its copy to the host
(SWESim_run_simulation). Inputs (blocks, field tables, initial rho) are in
HBM before the timed region. value = k-points * RK4 steps * K / wall time.

Printed JSON also carries: latency percentiles of the solves, J(t) relative
L2 error against the FP64 CPU oracle at full size, the roofline of the
propagation kernel (HIP events on its stream) and the CPU baseline (the
oracle's faithful real-space FFT restatement of the reference algorithm,
single-threaded like the reference, on a bounded sample of steps).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "hhg-solids_amd")]

METRIC = "k-point·RK4-steps/s; p50/p99 solve latency; J(t) rel-L2 err vs FP64"
UNIT = "k-point·RK4-steps/s"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}  # MI355X published vector peaks
ATTAINABLE_FP32 = 78.6  # non-packed FP32 VALU: one FMA per lane and instruction, the FP64 rate
PEAK_HBM_GBS = 8000.0
KERNELS = {1: "k_thread_per_k", 2: "k_pair2", 3: "k_elem_per_lane", 4: "k_pair2s", 6: "k_sep", 8: "k_bloch2", 9: "k_sep_lean",
           10: "k_bloch2w", 11: "k_bloch2s"}


def j_errors(t, J, Jref, lam):
    """J(t) relative L2 over the 3 x nt complex array and the worst relative
    deviation of the per-harmonic intensities up to the cutoff (the highest
    odd harmonic within 1e-6 of the strongest), per axis with a current."""
    from hhgamd import spectrum
    err = float(np.linalg.norm(J - Jref) / np.linalg.norm(Jref))
    worst = 0.0
    for a in range(3):
        if np.abs(Jref[:, a]).max() == 0:
            continue
        q, I = spectrum.harmonic_intensities(t, Jref[:, a], lam)
        _, I2 = spectrum.harmonic_intensities(t, J[:, a], lam)
        sel = q <= spectrum.cutoff_order(q, I)
        worst = max(worst, float(np.max(np.abs(I2[sel] / I[sel] - 1.0))))
    return err, worst


def workload_param(path_tb, nr, precision):
    return {"path_tb": path_tb, "nr1": nr, "nr2": nr, "nr3": 1, "tmax": 90.0, "dt": 0.02197,
            "intensity": 1e12, "lambda": 3000.0, "tmax_field": 80.0,
            "pol_vec": np.array([0.0, 1.0, 0.0]), "phi_vec": np.array([0.0, 0.0, 0.0]),
            "precision": precision}


def algorithmic_flops_per_kstep(nb, nR):
    """Generic (non-Hermitian) formulation: 8 (8 nb^3 + 4 nR nb^2 + 48 nb^2) --
    4 stages x 2 complex nb x nb products, the H and three rbc Fourier sums
    over nR blocks, the windowed J of three axes over 16 points (complex
    multiply-add = 8 flops). Reported beside the launched formulation's own
    count (SWESim_flops_per_kstep), which is what roofline.achieved uses."""
    return 8 * (8 * nb ** 3 + 4 * nR * nb ** 2 + 48 * nb ** 2)


def nonzero_blocks(H, r):
    return int(sum(1 for i in range(len(H)) if np.any(H[i] != 0) or np.any(r[:, i] != 0)))


def cpu_baseline(path_tb, nr, sample_steps):
    import oracle
    prm = workload_param(path_tb, nr, "fp64")
    prm["tmax"] = 0.02197 * sample_steps + 1e-9
    n = oracle.nt(prm)
    t0 = time.perf_counter()
    oracle.run(path_tb, prm, mode="fft")
    dt = time.perf_counter() - t0
    return {"value": nr * nr * n / dt, "unit": UNIT, "cores": 1, "kind": "port",
            "sample": f"oracle FFT-form restatement of the reference real-space loop (project-written FFT, "
                      f"not FFTW), {nr}x{nr} k-grid, {n} RK4 steps of the same workload, setup included, "
                      f"{dt:.1f} s"}


def c1_full(path_tb, threads):
    """The CPU-runnable configuration in full (30x30, 4096 RK4 steps): the GPU
    plain path, the oracle's FFT-form restatement of the reference loop on one
    core and its per-k form on `threads` host cores, each over the whole solve,
    and their J(t) agreement."""
    import oracle
    from hhgamd import lib
    prm = workload_param(path_tb, 30, "fp64")
    with lib.Solve(prm) as s:
        s.run()
        a = time.perf_counter()
        s.run()
        gpu_s = time.perf_counter() - a
        nk, nt = s.N, s.nt
        _, Jg = s.current()
    a = time.perf_counter()
    _, Jc = oracle.run(path_tb, prm, mode="fft")
    cpu_s = time.perf_counter() - a
    a = time.perf_counter()
    _, Jp = oracle.run(path_tb, prm, mode="perk", nthreads=threads)
    perk_s = time.perf_counter() - a
    return {"workload": "haldane-2band-30x30-nt4096-fp64", "gpu_value": nk * nt / gpu_s, "gpu_ms": gpu_s * 1e3,
            "cpu_fftform_value": nk * nt / cpu_s, "cpu_seconds": round(cpu_s, 2), "cpu_cores": 1,
            "cpu_perk_value": nk * nt / perk_s, "cpu_perk_seconds": round(perk_s, 2), "cpu_perk_cores": threads,
            "j_rel_l2_gpu_vs_cpu": float(np.linalg.norm(Jg - Jc) / np.linalg.norm(Jc)),
            "j_rel_l2_gpu_vs_cpu_perk": float(np.linalg.norm(Jg - Jp) / np.linalg.norm(Jp))}


def cpu_perk_threads(path_tb, nr, steps, threads):
    import oracle
    prm = workload_param(path_tb, nr, "fp64")
    prm["tmax"] = 0.02197 * steps + 1e-9
    n = oracle.nt(prm)
    t0 = time.perf_counter()
    oracle.run(path_tb, prm, mode="perk", nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": nr * nr * n / dt, "cores": threads, "steps": n, "seconds": round(dt, 2)}


def roofline_of(s, kms, nk, nt, precision, kernel_id, survey_nR):
    """Roofline of the propagation kernel: achieved = the model's minimal flops
    per k-point step (SWESim_min_flops_per_kstep, kernel-independent) x k-point
    steps / kernel time, against the published vector peak of the stage
    arithmetic; beside it the kernel's executed count and the survey's generic
    figures F and B."""
    F_min = s.min_flops_per_kstep()
    F_exec = s.flops_per_kstep()
    rate = nk * nt / (kms * 1e-3)
    F_gen = algorithmic_flops_per_kstep(s.nb, survey_nR)
    B_gen = 2 * (16 if precision == "fp64" else 8) * s.nb * s.nb
    achieved = F_min * rate / 1e12
    out = {"bound": "valu-fp64" if precision == "fp64" else "valu-fp32", "achieved": achieved,
           "peak": PEAK_TFLOPS[precision], "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[precision],
           "traffic": None, "kernel_ms": kms, "kernel": KERNELS.get(kernel_id),
           "flops_per_kstep_min": F_min, "flops_per_kstep_executed": F_exec,
           "executed_tflops": F_exec * rate / 1e12,
           "survey_flops_per_kstep": F_gen, "survey_equivalent_tflops": F_gen * rate / 1e12,
           "survey_bytes_per_kstep": B_gen, "survey_staged_state_GBs": B_gen * rate / 1e9}
    if precision != "fp64":
        out["attainable_peak"] = ATTAINABLE_FP32
        out["frac_attainable"] = achieved / ATTAINABLE_FP32
        out["note"] = ("published FP32 vector peak counts packed FP32 (2 FMAs per lane and instruction), which this "
                       "build does not emit; non-packed FP32 issues at the FP64 rate (attainable_peak)")
    return out


def traffic_of(workload):
    prof = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(prof):
        tr = json.load(open(prof)).get(workload)
        if tr:
            return tr["bytes_per_launch"]
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nr", type=int, default=None, help="k-grid side (default 120 for --mode precision, else 200)")
    ap.add_argument("--bands", type=int, default=8, help="precision: bands of the seeded model")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--cpu-steps", type=int, default=40, help="RK4 steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the full-size oracle J check")
    ap.add_argument("--no-legs", action="store_true",
                    help="plain mode: skip the precision (C3), batch (C4) and serving (C5) legs of the default line")
    ap.add_argument("--mode", default="plain", choices=["plain", "precision", "batch", "serve", "fftform", "fft2"],
                    help="plain: the FP64 200x200 solve (default) with the C3/C4/C5 legs; precision: the 8-band "
                         "120x120 solve with FP32 stages against FP64; batch: 64 mixed instances in one batched run; "
                         "serve: Poisson request streams through the continuous-batching server; fftform: the "
                         "reference's real-space algorithm on the GPU; fft2: its LDS radix 2-D FFT alone on "
                         "device-resident fields")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--fields", type=int, default=8, help="fft2: complex128 fields of nr x nr per transform")
    ap.add_argument("--autotune", action="store_true",
                    help="batch/serve/precision: tune every solve shape's launch configuration on its first run")
    ap.add_argument("--rates", default="20", help="serve: arrival rates (requests/s) to sweep")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests per rate")
    ap.add_argument("--check-requests", type=int, default=None,
                    help="serve: requests of the top rate also solved one by one on the plain FP64 path "
                         "(default: all of them, so the served and the plain rates cover the same requests)")
    ap.add_argument("--max-inflight", type=int, default=12, help="serve: solves on the device at once")
    ap.add_argument("--light-rate", type=float, default=2.0,
                    help="serve: arrival rate of the light-load stream (latency below saturation)")
    ap.add_argument("--light-requests", type=int, default=40,
                    help="serve: requests of the light-load stream (0: none)")
    ap.add_argument("--policy", default="least-work", choices=["least-work", "fifo"],
                    help="serve: admission order among prepared requests")
    args = ap.parse_args()
    if args.nr is None:
        args.nr = 120 if args.mode == "precision" else 200
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; runs on more than one GPU are not part of it "
              "(the k-grid would shard by slabs of k-rows, with one final reduce of J over the GPUs)", file=sys.stderr)
        sys.exit(2)

    from hhgamd import lib

    if args.autotune:
        lib.set_autotune(True)
    if args.mode == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.mode == "precision":
        print(json.dumps(bench_precision(args)), flush=True)
        return
    if args.mode == "fftform":
        return bench_fftform(args)
    if args.mode == "fft2":
        return bench_fft2(args)
    if args.mode == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    out = bench_plain(args)
    if not args.no_legs:
        # the other configurations at their full workloads, each in its own object
        leg = argparse.Namespace(**vars(args))
        leg.nr, leg.bands, leg.precision, leg.steps, leg.warmup = 120, 8, "fp32", 5, 1
        out["c3_precision"] = bench_precision(leg)
        print("c3 done", file=sys.stderr, flush=True)
        leg = argparse.Namespace(**vars(args))
        leg.precision, leg.steps, leg.warmup = "fp32", 2, 1
        out["c4_batch"] = bench_batch(leg)
        print("c4 done", file=sys.stderr, flush=True)
        leg = argparse.Namespace(**vars(args))
        leg.precision = "fp32"
        out["c5_serve"] = bench_serve(leg)
        print("c5 done", file=sys.stderr, flush=True)
        out["north_star_ratio"] = {
            "c4_batch_fp32_vs_plain_fp64": out["c4_batch"]["vs_plain_fp64"],
            "c5_served_fp32_vs_plain_fp64": out["c5_serve"]["vs_plain_fp64_at_top_rate"],
            "target": 5.0}
    print(json.dumps(out), flush=True)


def bench_plain(args):
    """The headline: the plain FP64 path on the 2-band Haldane-like model,
    200 x 200 k-grid, 4096 RK4 steps (a step = one solve)."""
    from hhgamd import lib, tbgen
    tmp = tempfile.mkdtemp(prefix="hhg_bench_")
    path_tb = os.path.join(tmp, "haldane_tb.dat")
    model = tbgen.haldane()
    tbgen.write_tb(path_tb, *model)
    prm = workload_param(path_tb, args.nr, args.precision)

    t_setup = time.perf_counter()
    s = lib.Solve(prm)
    setup_s = time.perf_counter() - t_setup
    nk, nt, nb = s.N, s.nt, s.nb
    kernel_id = s.kernel()
    for _ in range(args.warmup):
        s.run()
    lat, kern, Js = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        s.run()
        lat.append((time.perf_counter() - a) * 1e3)
        kern.append(s.kernel_ms())
    wall = time.perf_counter() - t0
    _, J = s.current()  # first solve after init (the first warmup run)
    roof = roofline_of(s, float(np.mean(kern)), nk, nt, args.precision, kernel_id,
                       nonzero_blocks(model[2], model[3]))
    # each run continues from the state the previous one left (src/swe.cpp:183-223
    # run twice): one more run from the state after the timed ones, checked below
    # against the oracle started from that same state
    rho_pre = s.rho_k() if not (args.no_cpu or args.no_check) else None
    s.run()
    _, J_cont = s.current_run(-1)
    s.close()
    value = nk * nt * args.steps / wall
    workload = f"haldane-2band-{args.nr}x{args.nr}-nt{nt}-{args.precision}"
    roof["traffic"] = traffic_of(workload)
    out = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": workload, "nk": nk, "nb": nb, "nt": nt,
                   "nR_nonzero": nonzero_blocks(model[2], model[3]), "precision": args.precision,
                   "setup_s": round(setup_s, 3)},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "roofline": roof,
    }
    if not args.no_cpu:  # CPU leg: the oracle, as baseline and as the FP64 checker of J(t)
        cores = cpu_cores()
        out["cpu_baseline"] = cpu_baseline(path_tb, args.nr, args.cpu_steps)
        out["c1_full"] = c1_full(path_tb, cores["used"])
        out["cpu_perk_threads"] = cpu_perk_threads(path_tb, args.nr, 512, cores["used"])
        out["cpu_perk_threads"].update(cores)
        if not args.no_check:
            import oracle
            _, Jref = oracle.run(path_tb, prm, mode="perk", nthreads=cores["used"])
            out["j_rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(J - Jref) / np.linalg.norm(Jref))
            _, Jc = oracle.run_from(path_tb, prm, rho_pre, nthreads=cores["used"])
            out["j_rel_l2_vs_fp64_oracle_after_timed_runs"] = float(np.linalg.norm(J_cont - Jc) / np.linalg.norm(Jc))
    return out


def cpu_cores():
    """Host cores for the multi-threaded CPU figure: the cores this process may
    run on, at most 16 (a one-GPU box's CPU share)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"used": max(1, min(16, avail)), "affinity": avail, "host_cpus": os.cpu_count()}


def bench_precision(args):
    """Seeded Wannier model (--bands, default 8; 5x5 R shell), nr x nr k-grid
    (default 120), I = 5e11, readme pulse otherwise: the mixed path (FP32
    commutator and RK4 stages, FP64 state, initial state and current) timed,
    its J(t) compared with the FP64 path of the same solve."""
    from hhgamd import lib, workloads
    tmp = tempfile.mkdtemp(prefix="hhg_bench_")
    prm = workloads.c3(tmp, precision="fp32", nb=args.bands, nr=args.nr)
    s = lib.Solve(prm)
    nk, nt, nb = s.N, s.nt, s.nb
    kpw = s.kpoints_per_wave()
    for _ in range(args.warmup):
        s.run()
    lat, kern = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        s.run()
        lat.append((time.perf_counter() - a) * 1e3)
        kern.append(s.kernel_ms())
    wall = time.perf_counter() - t0
    t, J32 = s.current()
    kernel_id = s.kernel()
    roof = roofline_of(s, float(np.mean(kern)), nk, nt, "fp32", kernel_id, 25)
    s.close()
    with lib.Solve(dict(prm, precision="fp64")) as s64:
        s64.run()
        t64 = time.perf_counter()
        s64.run()
        fp64_s = time.perf_counter() - t64
        _, J64 = s64.current()
        fp64_kms = s64.kernel_ms()
        roof64 = roofline_of(s64, fp64_kms, nk, nt, "fp64", s64.kernel(), 25)
    err, worst = j_errors(t, J32, J64, prm["lambda"])
    cpu = None
    if not args.no_cpu:  # the oracle's per-k FP64 form on the host cores, a bounded sample of steps
        import oracle
        cores = cpu_cores()
        steps = 64
        p64 = dict(prm, precision="fp64", tmax=prm["dt"] * steps + 1e-9)
        a = time.perf_counter()
        oracle.run(prm["path_tb"], p64, mode="perk", nthreads=cores["used"])
        cs = time.perf_counter() - a
        cpu = {"value": nk * oracle.nt(p64) / cs, "cores": cores["used"], "kind": "port",
               "sample": f"oracle per-k FP64 form, {oracle.nt(p64)} RK4 steps of the same solve, {cs:.1f} s",
               **cores}
    workload = f"synthetic-{nb}band-{prm['nr1']}x{prm['nr2']}-nt{nt}-fp32"
    roof["traffic"] = traffic_of(workload)
    return {"metric": METRIC, "value": nk * nt * args.steps / wall, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": workload, "nk": nk, "nb": nb,
                       "nt": nt, "precision": "fp32 stages, fp64 state/current", "kpoints_per_lane_group": kpw},
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
            "j_rel_l2_vs_fp64": err, "worst_harmonic_dev_vs_fp64": worst,
            "fp64_value": nk * nt / fp64_s, "fp64_kernel_ms": fp64_kms, "fp32_vs_fp64": fp64_s / (wall / args.steps),
            "cpu_perk_threads": cpu, "roofline": roof, "roofline_fp64": roof64}


def bench_fftform(args):
    """The reference's real-space algorithm on the GPU (kernel variant 2:
    rho(R), LDS-resident radix FFTs of 3 nb^2 fields per RK4 stage, one step
    replayed as a HIP graph) on the plain workload, against the per-k path.
    Streamed complex128 fields per k-point and RK4 step, counting each field
    read or written once per launch:
      fused stages (nb <= 4 and a column of all 2 nb^2 fields in LDS): per
      stage, forming + rows 6 nb^2 + 1 read (+ nb^2 for K_prev after the first
      stage) and 2 nb^2 written, columns + commutator 2 nb^2 + nb^2, inverse
      rows nb^2 + nb^2; phase 1, update 6 nb^2 + 1 -> 61 nb^2 + 6 fields;
      unfused: per stage forming 8, forward FFT rows + columns of 2 nb^2 fields
      8, commutator 3, inverse FFT 4 (x nb^2), update 6 -> 98 nb^2 fields."""
    from hhgamd import lib, tbgen
    tmp = tempfile.mkdtemp(prefix="hhg_bench_")
    path_tb = os.path.join(tmp, "haldane_tb.dat")
    tbgen.write_tb(path_tb, *tbgen.haldane())
    prm = workload_param(path_tb, args.nr, "fp64")
    s = lib.Solve(dict(prm, kernel="fftform"))
    nk, nt, nb = s.N, s.nt, s.nb
    for _ in range(args.warmup):
        s.run()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        s.run()
        lat.append((time.perf_counter() - a) * 1e3)
    wall = time.perf_counter() - t0
    _, Jf = s.current()
    s.close()
    with lib.Solve(prm) as sk:
        sk.run()
        _, Jk = sk.current()
    fused = nb <= 4 and (2 * nb * nb + 2) * args.nr * 16 <= 160 * 1024
    bytes_per_kstep = 16 * (61 * nb * nb + 6) if fused else 98 * 16 * nb * nb
    ms = wall / args.steps * 1e3
    gbs = bytes_per_kstep * nk * nt / (ms * 1e-3) / 1e9
    out = {"metric": METRIC, "value": nk * nt * args.steps / wall, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"haldane-2band-{args.nr}x{args.nr}-nt{nt}-fp64-fftform", "nk": nk, "nb": nb,
                      "nt": nt, "kernel": "FFT form (reference algorithm), graph per RK4 step",
                      "fused_stages": fused},
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
           "j_rel_l2_vs_per_k": float(np.linalg.norm(Jf - Jk) / np.linalg.norm(Jk)),
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": gbs / PEAK_HBM_GBS, "traffic": None,
                        "bytes_per_kstep": bytes_per_kstep, "note": "whole-step streamed bytes over solve time"}}
    print(json.dumps(out), flush=True)


def bench_fft2(args):
    """The LDS-resident radix 2-D FFT of the FFT form alone, on device-resident
    data (hhg_fft2_plan_exec on a hipMalloc'd array): by default 2 nb^2 = 8
    complex128 fields of nr x nr, the forward transform of one RK4 stage of the
    plain workload (--fields for larger batches).
    Algorithmic bytes per 2-D transform: two passes (rows, then columns), each
    reading and writing every field once = 2 * 2 * fields * nr^2 * 16 B; one pass
    is one k_fft_lines launch."""
    from hhgamd import lib
    nch, n = args.fields, args.nr
    rng = np.random.default_rng(11)
    a = rng.normal(size=(nch, n, n)) + 1j * rng.normal(size=(nch, n, n))
    with lib.FFT2Plan(nch, n, n) as plan, lib.DeviceArray((nch, n, n)) as x:
        x.upload(a)
        plan.exec(x.address, -1)
        ref = np.fft.fft2(a)
        err = float(np.abs(x.download() - ref).max() / np.abs(ref).max())
        plan.exec(x.address, -1, max(args.warmup, 1))
        ms = plan.exec(x.address, -1, args.steps)
    bytes_per_fft = 2 * 2 * nch * n * n * 16
    gbs = bytes_per_fft * args.steps / (ms * 1e-3) / 1e9
    flops = 5.0 * nch * n * n * np.log2(n * n)
    out = {"metric": "LDS radix 2-D FFT throughput (complex128, in place, device-resident)", "value": gbs,
           "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c128", "data": "synthetic",
           "config": {"workload": f"fft2-{nch}x{n}x{n}-c128", "fields": nch, "n1": n, "n2": n,
                      "passes_per_fft": 2},
           "rel_err_vs_numpy": err,
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": gbs / PEAK_HBM_GBS, "traffic": None, "bytes_per_fft": bytes_per_fft,
                        "gflops_5nlogn": flops * args.steps / (ms * 1e-3) / 1e9}}
    prof = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(prof):
        tr = json.load(open(prof)).get(out["config"]["workload"])
        if tr:  # both passes of one 2-D transform
            out["roofline"]["traffic"] = tr["bytes_per_launch"]
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """64 seeded instances (pol angle, intensity, wavelength, nb in {2,4,8},
    uneven 5-smooth k-grids) propagated by one batched run per step. Also
    timed: the same instances one by one in the same precision, and the plain
    FP64 path one by one (the reference point of the mixed/batched form);
    every member's J(t) is compared with its FP64 solve."""
    from hhgamd import lib, workloads
    tmp = tempfile.mkdtemp(prefix="hhg_bench_")
    params = workloads.c4(tmp, n=args.instances, precision=args.precision)
    solves = [lib.Solve(p) for p in params]
    total = sum(s.N * s.nt for s in solves)
    batch = lib.Batch(solves)
    for _ in range(args.warmup):
        batch.run()
    lat, dev = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        dev.append(batch.run())
        lat.append((time.perf_counter() - a) * 1e3)
    wall = time.perf_counter() - t0
    batch.close()
    # the same instances one by one (one pass)
    t1 = time.perf_counter()
    for s in solves:
        s.run()
    seq = time.perf_counter() - t1
    kinds = {}  # launch configurations of the batch members: band count / kernel / waves / k-points per group
    for s, p in zip(solves, params):
        L = s.batch_launch()
        key = f"nb{s.nb}/k{L['kernel']}" + (f"/w{L['b']}/kpw{L['kpw']}" if L["kernel"] in (6, 9) else "")
        kinds[key] = kinds.get(key, 0) + 1
    # plain FP64 path, one solve at a time, same instances; J(t) of each for the error check
    errs, worsts = [], []
    plain_s = 0.0
    for s, p in zip(solves, params):
        t, Jb = s.current()
        with lib.Solve(dict(p, precision="fp64")) as s64:
            a = time.perf_counter()
            s64.run()
            plain_s += time.perf_counter() - a
            _, J64 = s64.current()
        if args.precision != "fp64":
            e, w = j_errors(t, Jb, J64, p["lambda"])
            errs.append(e)
            worsts.append(w)
    for s in solves:
        s.close()
    value = total * args.steps / wall
    out = {"metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
           "data": "synthetic",
           "config": {"workload": f"batch-{len(params)}-mixed-nb248-{args.precision}", "instances": len(params),
                      "kpoint_steps_per_pass": total, "kernels": kinds, "autotune": bool(args.autotune)},
           "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
           "device_ms_per_pass": float(np.mean(dev)),
           "one_by_one_value": total / seq, "batch_speedup_vs_one_by_one": seq / (wall / args.steps),
           "plain_fp64_one_by_one_value": total / plain_s, "vs_plain_fp64": value * plain_s / total}
    if errs:
        out["j_rel_l2_vs_fp64_max"] = max(errs)
        out["worst_harmonic_dev_vs_fp64"] = max(worsts)
    return out


def bench_serve(args):
    """Poisson streams of mixed solves (nb in {2,4,8,16}, 5-smooth grids in
    [48, 240], tmax U[60, 150]) through the server at each arrival rate
    (default: --requests 200 at 20/s, a sustained stream past saturation),
    then --light-requests of the same generator at --light-rate (latency
    below saturation);
    reports p50/p99 latency (arrival to J available) and k-point*RK4-steps/s.
    Every rate replays the same requests (one seed; the exponential gaps scale
    with 1/rate). The requests of the top rate (the first --check-requests
    when given) are also solved one by one on the plain FP64 path: their
    throughput is the reference point of the served rate and their currents
    check the served FP32 results. The C4 instance mix is also streamed through the server at
    100 requests/s, the like-for-like comparison with the batched rate."""
    from hhgamd import lib, serve, workloads
    tmp = tempfile.mkdtemp(prefix="hhg_bench_")
    rates = [float(r) for r in args.rates.split(",")]
    sweep = []
    recs = []
    with serve.Server(hbm_budget_gb=128.0, max_inflight=args.max_inflight, host_threads=8, policy=args.policy) as srv:
        warm = workloads.c5(tmp, 2, rate=100.0, seed=99, precision=args.precision)
        serve.run_stream(srv, [(0.0, p) for _, p in warm])
        for ri, rate in enumerate(rates):
            arrivals = workloads.c5(tmp, args.requests, rate=rate, seed=50, precision=args.precision)
            recs, summ = serve.run_stream(srv, arrivals, keep_results=ri == len(rates) - 1,
                                          keep_first=args.check_requests)
            summ["rate_per_s"] = rate
            sweep.append(summ)
            print(json.dumps({"serve_rate": rate, **summ}), file=sys.stderr, flush=True)
        light = None
        if args.light_requests > 0:  # the same generator's first requests, spaced out
            la = workloads.c5(tmp, args.light_requests, rate=args.light_rate, seed=50, precision=args.precision)
            _, light = serve.run_stream(srv, la)
            light["rate_per_s"] = args.light_rate
            print(json.dumps({"serve_light_rate": args.light_rate, **light}), file=sys.stderr, flush=True)
        c4 = workloads.c4(tmp, n=64, precision=args.precision)
        _, c4summ = serve.run_stream(srv, [(i / 100.0, p) for i, p in enumerate(c4)])
    top = len(rates) - 1
    plain_s, plain_ksteps = 0.0, 0
    errs = []
    checked = workloads.c5(tmp, args.requests, rate=rates[top], seed=50, precision="fp64")
    if args.check_requests is not None:
        checked = checked[:args.check_requests]
    for rec, (_, p) in zip(recs, checked):
        with lib.Solve(p) as s:
            a = time.perf_counter()
            s.run()
            plain_s += time.perf_counter() - a
            plain_ksteps += s.N * s.nt
            if "J" in rec:
                t64, J64 = s.current()
                errs.append(j_errors(t64, rec["J"], J64, p["lambda"]))
    best = max(sweep, key=lambda x: x["ksteps_per_s"])
    plain = plain_ksteps / plain_s if plain_s > 0 else None
    return {"metric": METRIC, "value": best["ksteps_per_s"], "unit": UNIT, "n_gpus": 1, "steps": args.requests,
            "warmup": 2, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
            "config": {"workload": f"serve-poisson-nb2-16-{args.precision}", "requests_per_rate": args.requests,
                       "rates_per_s": rates, "policy": args.policy},
            "latency_ms": {"p50": best["p50_latency_s"] * 1e3, "p99": best["p99_latency_s"] * 1e3},
            "plain_fp64_one_by_one_value": plain,
            "plain_fp64_sample": f"{len(checked)} of the {args.requests} requests of the top rate, one by one",
            "vs_plain_fp64_at_top_rate": sweep[top]["ksteps_per_s"] / plain if plain else None,
            "served_vs_fp64": {"results": len(errs),
                               "max_j_rel_l2": max((e[0] for e in errs), default=None),
                               "worst_harmonic_dev": max((e[1] for e in errs), default=None)},
            "light_load": light,
            "c4_mix_served": {"value": c4summ["ksteps_per_s"], "requests": c4summ["requests"],
                              "rate_per_s": 100.0, "p50_latency_s": c4summ["p50_latency_s"],
                              "p99_latency_s": c4summ["p99_latency_s"]},
            "sweep": sweep}


if __name__ == "__main__":
    main()
