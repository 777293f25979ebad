#!/usr/bin/env python3
"""Benchmark of the BLS12-381 prover hot path on one MI355X.

Default workload: G1 MSM with 2^20 points (headline metric "G1 MSM Mpts/s
@2^20").  A step is one call of the resident-context batch entry point
(midnight_sppark_bls12_381_g1_msm_context_device_batch) over
--msms-per-step MSMs of 2^20 uniform 255-bit scalars against one resident
point set [t_i]G.  Points and scalars are synthetic, generated on the device
before timing, so the timed region starts with every input in HBM; each
call returns its Jacobian results to the host before the next starts.

The default run also reports, in the same JSON line:
- "one_shot": the same 2^20 MSM through the one-shot entry point (host
  inputs uploaded by the call);
- "serving": the proof server on the serving mix (k in {12, 14, 16, 18} at
  50/30/15/5 %, SRS 2^20 resident), batched and one proof at a time on the
  same Poisson streams at 4x and at 0.9x the one-at-a-time capacity, with
  p50/p99 latency and the CPU reference prover's proofs/s on the mix;
- "cpu_baseline": msm_best (batch-affine buckets) on the same 2^20 MSM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "midnight-ledger-cuda_amd"))

R_MOD = 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001
R_LIMBS = np.array([(R_MOD >> (64 * i)) & (2**64 - 1) for i in range(4)], dtype=np.uint64)

# Integer-VALU peak: v_mad_u64_u32 at 64 lanes/clk/CU x 256 CUs x 2.4 GHz
# (the published CU count and peak clock).  `roofline.peak` and `frac` use
# this figure; the rate measured live on the box by midnight_hip_valu_peak
# (below it) is reported beside it as `peak_measured`.
PEAK_SPEC_MAC_PER_S = 256 * 64 * 2.4e9
HBM_PEAK = 8.0e12
# algorithmic Montgomery-product cost of one Fp product (12 limbs, CIOS)
MACS_PER_FP_MUL = 300
# algorithmic MSM work: min over c of N*ceil(256/c)*(8M+2S) + ceil(256/c)*2^(c-1)*2*(12M+2S)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_calibration(H, mib):
    """Device-to-device copies of `mib` MiB by a plain copy kernel
    (k_copy16, one untimed and one timed launch), run before the timed
    region so that hardware-counter runs can scale the FETCH_SIZE /
    WRITE_SIZE they read for the other kernels."""
    rate = H.copy_bandwidth(mib << 20)
    log("copy calibration: %d MiB, %.0f GB/s read+write" % (mib, rate / 1e9))


def profiled_traffic(name):
    """HBM bytes per launch of the dominant kernel, from the hardware-counter
    summary committed under profiles/ (FETCH_SIZE + WRITE_SIZE, each scaled
    by the counter's reading for a copy of known size); None if absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def msm_algorithmic_macs(n):
    best = None
    for c in range(1, 25):
        w = -(-256 // c)
        fp = n * w * 10 + w * (1 << (c - 1)) * 2 * 14
        best = fp if best is None or fp < best else best
    return best * MACS_PER_FP_MUL


def random_canonical(n, rng):
    """uniform values in [0, r): 255-bit draws, rejection of >= r"""
    out = np.empty((n, 4), dtype=np.uint64)
    filled = 0
    while filled < n:
        m = int((n - filled) * 1.1) + 16
        x = rng.integers(0, 2**64, size=(m, 4), dtype=np.uint64, endpoint=False)
        x[:, 3] &= np.uint64((1 << 63) - 1)
        # lexicographic x < r on limbs 3..0
        lt = np.zeros(m, dtype=bool)
        eq = np.ones(m, dtype=bool)
        for i in (3, 2, 1, 0):
            lt |= eq & (x[:, i] < R_LIMBS[i])
            eq &= x[:, i] == R_LIMBS[i]
        x = x[lt]
        take = min(n - filled, x.shape[0])
        out[filled:filled + take] = x[:take]
        filled += take
    return out


def skew(vals, rng):
    """70% in [0, 2^16), 20% in {0, 1}, 10% unchanged (uniform)"""
    u = rng.random(vals.shape[0])
    small = u < 0.7
    zo = (u >= 0.7) & (u < 0.9)
    vals[small] = 0
    vals[small, 0] = rng.integers(0, 1 << 16, size=int(small.sum()), dtype=np.uint64)
    vals[zo] = 0
    vals[zo, 0] = rng.integers(0, 2, size=int(zo.sum()), dtype=np.uint64)
    return vals


def device_fr(H, canonical):
    dv = H.DeviceVec.upload(canonical)
    H.to_montgomery(dv)
    return dv


def cpu_threads():
    """host threads for the CPU baselines: all cores, at most 16 (one GPU's
    share of its host)"""
    return min(16, os.cpu_count() or 1)


def cpu_baseline_msm(H, ctx, dv_scalars, log_n, seconds):
    """msm_best (the reference's CPU MSM above 2^19: Booth windows of
    ceil(ln n) bits with the batch-affine bucket schedule, restated in
    oracle/msm.c) on the workload's first MSM at full size (up to 2^20
    points), repeated for about `seconds` on all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sample_log = min(log_n, 20)
    n = 1 << sample_log
    pts = ctx.points(n)
    sc = dv_scalars.download()[:n]
    threads = cpu_threads()
    oracle.msm(pts[:16384], sc[:16384], threads=threads, best=True)     # warm-up (threads)
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.msm(pts, sc, threads=threads, best=True)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n * reps / dt / 1e6, 4), "unit": "Mpts/s", "cores": threads, "kind": "port",
            "sample": "%d x MSM of 2^%d points (the workload's first MSM), msm_best with batch-affine buckets "
                      "restated in C, %d threads, %.1f s" % (reps, sample_log, threads, dt)}


def bench_msm(args, H):
    log_n = args.log_n
    n = 1 << log_n
    m = args.msms_per_step
    rng = np.random.default_rng(args.seed)
    log("generating %d points and %d x %d scalars on the device" % (n, m, n))
    t_dv = device_fr(H, random_canonical(n, rng))
    ctx = H.MsmContext.from_multiples(t_dv)
    t_dv.free()
    sc = random_canonical(n * m, rng)
    if args.skewed:
        sc = skew(sc, rng)
    dv = device_fr(H, sc)
    del sc
    for _ in range(args.warmup):
        ctx.device_batch(dv, 0, n, m)
    if args.pmc_copy_mib:
        copy_calibration(H, args.pmc_copy_mib)
    peak_measured, _, _ = H.valu_peak()
    peak = PEAK_SPEC_MAC_PER_S
    H.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.device_batch(dv, 0, n, m)
    dt = time.perf_counter() - t0
    acc_ms, acc_cnt = H.profile_read("msm_accumulate")
    tot_ms, tot_cnt = H.profile_read("msm_total")
    H.profile_enable(False)
    entries, window = ctx.last_stats()
    tuned_w, tuned_l, tuned_ms = ctx.tuning()
    ms_step = dt / args.steps * 1e3
    value = n * m * args.steps / dt / 1e6
    # dominant kernel: bucket accumulation, 8M+2S Fp products per entry
    acc_avg_s = acc_ms / max(acc_cnt, 1) * 1e-3
    acc_macs = entries * 10 * MACS_PER_FP_MUL
    achieved = acc_macs / acc_avg_s if acc_avg_s > 0 else 0.0
    msm_macs = msm_algorithmic_macs(n) * m
    res = {
        "metric": "G1 MSM Mpts/s @2^%d" % log_n,
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: points [t_i]G and %s scalars, generated on device" % ("skewed" if args.skewed else "uniform"),
        "config": {"workload": "G1 MSM, 2^%d resident points, %d MSMs per step (%s scalars)"
                               % (log_n, m, "skewed" if args.skewed else "uniform 255-bit"),
                   "points": n, "msms_per_step": m, "window_bits": window,
                   "autotuned": {"table_window": tuned_w, "chunk_len": tuned_l, "tuning_ms": round(tuned_ms, 3)}},
        "roofline": {"bound": "valu", "kernel": "msm_accumulate",
                     "achieved": round(achieved / 1e12, 3), "peak": round(peak / 1e12, 3),
                     "peak_measured": round(peak_measured / 1e12, 3),
                     "unit": "TMAC/s", "frac": round(achieved / peak, 4),
                     "traffic": profiled_traffic("msm_accumulate_2p20_traffic.md")
                     if (log_n, m, args.skewed) == (20, 4, False) else None},
        "msm_valu_frac": round(msm_macs / (dt / args.steps) / peak, 4),
        "accumulate_ms_avg": round(acc_avg_s * 1e3, 3),
        "engine_ms_avg": round(tot_ms / max(tot_cnt, 1), 3),
    }
    if log_n <= 22:
        # the same MSM through the one-shot entry point (host points and
        # scalars, uploaded by the call; no resident context or table)
        pts = ctx.points(n)
        sc_h = dv.download()[:n]
        H.g1_msm(pts[:4096], sc_h[:4096])
        reps, t0 = 3, time.perf_counter()
        for _ in range(reps):
            H.g1_msm(pts, sc_h)
        dt1 = (time.perf_counter() - t0) / reps
        res["one_shot"] = {"mpts_per_s": round(n / dt1 / 1e6, 3), "ms_per_msm": round(dt1 * 1e3, 3),
                           "entry": "midnight_sppark_bls12_381_g1_msm (host inputs, PCIe upload included)"}
        del pts, sc_h
    if args.cpu_seconds > 0:
        log("timing the CPU baseline")
        res["cpu_baseline"] = cpu_baseline_msm(H, ctx, dv, log_n, args.cpu_seconds)
    ctx.free()
    dv.free()
    if args.serving:
        res["serving"] = serving_report(args, H, sorted({args.load, 0.9}, reverse=True))
    return res


def bench_ntt(args, H):
    log_n = args.log_n
    n = 1 << log_n
    cnt = args.msms_per_step
    rng = np.random.default_rng(args.seed)
    dv = device_fr(H, random_canonical(n * cnt, rng))
    one = np.array([0x00000001fffffffe, 0x5884b7fa00034802, 0x998c4fefecbc4ff5, 0x1824b159acc5056f], dtype=np.uint64)
    for _ in range(args.warmup):
        dv.ntt_zeta_batch(log_n, cnt, False, one, one, one)
    t0 = time.perf_counter()
    for i in range(args.steps):
        dv.ntt_zeta_batch(log_n, cnt, bool(i & 1), one, one, one)
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    # algorithmic work: (n/2) log2 n Fr products of 2*8^2 + 8 = 136 MACs; bytes: one read and one write
    bytes_alg = 64 * n * cnt
    macs = 68 * n * log_n * cnt
    peak_measured, _, _ = H.valu_peak()
    peak = PEAK_SPEC_MAC_PER_S
    t_step = dt / args.steps
    res = {
        "metric": "Fr NTT 2^%d elements/s" % log_n, "value": round(n * cnt * args.steps / dt / 1e6, 3),
        "unit": "Melem/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic uniform Fr",
        "config": {"workload": "resident Fr NTT 2^%d, %d polys per step" % (log_n, cnt)},
        "roofline": {"bound": "valu", "achieved": round(macs / t_step / 1e12, 3), "peak": round(peak / 1e12, 3),
                     "peak_measured": round(peak_measured / 1e12, 3),
                     "unit": "TMAC/s", "frac": round(macs / t_step / peak, 4), "traffic": None},
        "hbm": {"achieved": round(bytes_alg / t_step / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(bytes_alg / t_step / HBM_PEAK, 4)},
    }
    if args.cpu_seconds > 0:
        log("ntt: timing the CPU reference FFT")
        res["cpu_baseline"] = cpu_baseline_ntt(dv.download()[:n], log_n, args.cpu_seconds)
    dv.free()
    return res


def cpu_baseline_ntt(values, log_n, seconds):
    """The reference's best_fft (radix-2 with the parallel split of
    fft.rs:30-112, restated in oracle/ntt.c) on the workload's first
    polynomial, repeated for about `seconds` on all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = cpu_threads()
    oracle.fft(values[:1024], 10, threads=threads)   # warm-up (threads)
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.fft(values, log_n, threads=threads)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round((len(values) * reps) / dt / 1e6, 3), "unit": "Melem/s", "cores": threads, "kind": "port",
            "sample": "%d x FFT of 2^%d values (the workload's first polynomial), best_fft restated in C, "
                      "%d threads, %.1f s" % (reps, log_n, threads, dt)}


def bench_batch(args, H):
    """256 independent instances, sizes 2^k with k log-uniform in [10, 18]
    (k = floor(10 + 9 u), u from a per-instance u64 seed): each does an
    iNTT of 2^k values and an MSM of 2^k points over a shared 2^18 point set.
    A step is one run of a batch plan: the ragged iNTT (one launch per pass
    for all instances) and one MSM batch, captured once in a HIP graph and
    replayed, with the MSM configuration autotuned when the plan is made.
    The same work without the graph (eager launches) and one instance at a
    time is timed beside it."""
    count = 256
    ks = []
    for j in range(count):
        u = np.random.default_rng(args.seed * 1000003 + j).random()
        ks.append(min(18, max(10, int(10 + 9 * u))))
    order = sorted(range(count), key=lambda j: ks[j])
    offs, total = {}, 0
    for j in order:
        offs[j] = total
        total += 1 << ks[j]
    rng = np.random.default_rng(args.seed)
    t_dv = device_fr(H, random_canonical(1 << 18, rng))
    ctx = H.MsmContext.from_multiples(t_dv)
    t_dv.free()
    base = random_canonical(total, rng)
    dv = device_fr(H, base)
    o_list = [offs[j] for j in order]
    k_list = [ks[j] for j in order]
    n_list = [1 << k for k in k_list]

    t0 = time.perf_counter()
    plan = H.BatchPlan(ctx, dv, o_list, k_list, H.BatchPlan.GRAPH | H.BatchPlan.AUTOTUNE)
    plan_s = time.perf_counter() - t0
    info = plan.info()
    log("batch plan: %s (made in %.2f s)" % (info, plan_s))

    def batched():
        plan.run()

    def eager():
        dv.ntt_multi(o_list, k_list, True)
        ctx.device_multi(dv, o_list, n_list)

    def singles():
        for o, k in zip(o_list, k_list):
            dv.ntt_multi([o], [k], True)
            ctx.device_multi(dv, [o], [1 << k])

    for _ in range(args.warmup):
        batched()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
    dt = (time.perf_counter() - t0) / args.steps
    eager()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eager()
    dt_eager = (time.perf_counter() - t0) / args.steps
    singles()
    t0 = time.perf_counter()
    singles()
    dt1 = time.perf_counter() - t0
    plan.free()
    res = {
        "metric": "batched MSM+iNTT instances/s (256 instances, k in [10,18])",
        "value": round(count / dt, 2), "unit": "instances/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic uniform Fr, points [t_i]G",
        "config": {"workload": "256 x (iNTT 2^k + MSM 2^k), k = floor(10 + 9u) per instance seed, shared 2^18 points",
                   "total_points": total, "plan": info},
        "eager_batched_ms": round(dt_eager * 1e3, 3),
        "one_at_a_time_ms": round(dt1 * 1e3, 3),
        "batched_speedup": round(dt1 / dt, 3),
    }
    if args.cpu_seconds > 0:
        log("batch: timing the CPU reference path on a sample of the instances")
        res["cpu_baseline"] = cpu_baseline_batch(ctx, base, o_list, k_list, args.cpu_seconds)
    ctx.free()
    dv.free()
    return res


def cpu_baseline_batch(ctx, base, o_list, k_list, seconds):
    """The reference's CPU path per instance (best_fft inverse, then the
    CPU MSM: Pippenger below 2^19 points, msm_best from 2^19) on the
    instances in the batch's order until about `seconds` have passed; the
    batch's instances/s is extrapolated by points (sample = the instances
    done)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = cpu_threads()
    pts = ctx.points(1 << max(k_list))
    done_pts, done, t0 = 0, 0, time.perf_counter()
    for o, k in zip(o_list, k_list):
        n = 1 << k
        coeffs = oracle.ifft(base[o:o + n], k, threads=threads)
        oracle.msm(pts[:n], coeffs, threads=threads, best=k >= 19)
        done_pts += n
        done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    total_pts = sum(1 << k for k in k_list)
    est = dt * total_pts / done_pts
    return {"value": round(len(k_list) / est, 3), "unit": "instances/s", "cores": threads, "kind": "port",
            "sample": "%d of %d instances (%d of %d points, the smallest first) in %.1f s on %d threads: iFFT "
                      "(best_fft) + CPU MSM restated in C; the whole batch extrapolated by points to %.1f s"
                      % (done, len(k_list), done_pts, total_pts, dt, threads, est)}


def bench_kzg(args, H):
    """Plain KZG commitment of 2^log_n Lagrange values (default 2^16): the
    values are uploaded from the host, inverse-NTT'd to coefficients and
    committed with one MSM against the SRS [tau^i]G, resident on the device
    (tau from the seed).  A step is one commitment; the value is
    commitments per second with the values starting in host memory, as the
    prover's commit_lagrange receives them."""
    log_n = args.log_n if args.log_n != 20 else 16
    n = 1 << log_n
    rng = np.random.default_rng(args.seed)
    tau = random_canonical(1, rng)
    tau_dv = device_fr(H, tau)
    tau_m = tau_dv.download()[0]
    tau_dv.free()
    powers = H.fr_shifted_prefix_product(np.tile(tau_m, (n, 1)), MONT_ONE)
    pw = H.DeviceVec.upload(powers)
    ctx = H.MsmContext.from_multiples(pw)
    pw.free()
    evals = random_canonical(n, rng)      # any value < r is a valid Montgomery element

    def commit():
        dv = H.DeviceVec.upload(evals)
        dv.ntt_zeta(log_n, True, MONT_ONE, MONT_ONE, MONT_ONE)
        out = ctx.device_batch(dv, 0, n, 1)
        dv.free()
        return out

    for _ in range(args.warmup):
        commit()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        commit()
        lat.append(time.perf_counter() - t1)
    dt = (time.perf_counter() - t0) / args.steps
    lat.sort()
    res = {
        "metric": "KZG commit 2^%d commitments/s" % log_n, "value": round(1.0 / dt, 2), "unit": "commits/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic uniform Fr values on the host; SRS [tau^i]G built on the device",
        "config": {"workload": "KZG commit: upload 2^%d Lagrange values, iNTT, MSM against [tau^i]G" % log_n},
        "p50_ms": round(lat[len(lat) // 2] * 1e3, 3), "p99_ms": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))] * 1e3, 3),
    }
    if args.cpu_seconds > 0:
        log("kzg: timing the CPU reference commit")
        res["cpu_baseline"] = cpu_baseline_kzg(ctx.points(n), evals, log_n, args.cpu_seconds)
    ctx.free()
    return res


def cpu_baseline_kzg(pts, evals, log_n, seconds):
    """The reference's CPU commitment of Lagrange values: best_fft inverse
    to coefficients, then the CPU MSM (blst-style Pippenger below 2^19
    points, msm_best from 2^19), restated in C, on the same values,
    repeated for about `seconds` on all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = cpu_threads()
    reps, t0 = 0, time.perf_counter()
    while True:
        coeffs = oracle.ifft(evals, log_n, threads=threads)
        oracle.msm(pts, coeffs, threads=threads, best=log_n >= 19)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(reps / dt, 3), "unit": "commits/s", "cores": threads, "kind": "port",
            "sample": "%d commitments of the workload's 2^%d values (iFFT + MSM restated in C), %d threads, %.1f s"
                      % (reps, log_n, threads, dt)}


MONT_ONE = np.array([0x00000001fffffffe, 0x5884b7fa00034802, 0x998c4fefecbc4ff5, 0x1824b159acc5056f],
                    dtype=np.uint64)


def synthetic_circuit(H, k, seed):
    """One instance of the benchmark circuit family at 2^k rows: advice
    a, b, c uniform, d = a*b + c on the usable rows (products on the
    device), e copying a quarter of the usable rows of a (copy cycles of
    two), fixed q = 1 on usable rows, s1, s2 uniform there.  Returns (fixed,
    advice) as Montgomery arrays and the permutation mapping."""
    n = 1 << k
    u = n - 6
    rng = np.random.default_rng(seed)
    dv = device_fr(H, random_canonical(7 * n, rng))
    vals = dv.download()
    dv.free()
    a, b, c, d, e, s1, s2 = (vals[i * n:(i + 1) * n].copy() for i in range(7))
    ab = H.fr_mul_assign_repeating(a[:u], b[:u])                    # a_i * b_i
    d[:u] = H.fr_poly_lincomb(np.concatenate([ab, c[:u]]), u, np.tile(MONT_ONE, (2, 1)))
    q = np.zeros((n, 4), dtype=np.uint64)
    q[:u] = MONT_ONE
    s1[u:] = 0
    s2[u:] = 0
    mapping = np.zeros((8, n, 2), dtype=np.uint32)
    mapping[:, :, 0] = np.arange(8, dtype=np.uint32)[:, None]
    mapping[:, :, 1] = np.arange(n, dtype=np.uint32)[None, :]
    rows = rng.permutation(u)
    pairs = u // 4
    ei, aj = rows[:pairs], rows[pairs:2 * pairs]
    e[ei] = a[aj]
    mapping[4, ei, 0], mapping[4, ei, 1] = 0, aj
    mapping[0, aj, 0], mapping[0, aj, 1] = 4, ei
    return np.concatenate([q, s1, s2]), np.concatenate([a, b, c, d, e]), mapping.reshape(-1)


def synthetic_argument_circuit(H, k, seed, m=6):
    """The benchmark circuit plus every argument of the prover: advice f, g
    holding random rows of fixed tables t1, t2 (lookup (f, g) in (t1, t2)),
    lookup q * e in s1 (e's copied cells' values placed in s1 at their rows,
    its other cells taking s1 values), a trashcan with selector qt on odd
    usable rows over (d - ab - c, e (d - ab - c)), and an instance column p
    of m values equal to a's first rows, tied by the gate qp (a - p).
    Returns (fixed, advice, instance, mapping)."""
    n = 1 << k
    u = n - 6
    fixed, adv, mapping = synthetic_circuit(H, k, seed)
    q, s1, s2 = (fixed[i * n:(i + 1) * n].copy() for i in range(3))
    a, b, c, d, e = (adv[i * n:(i + 1) * n].copy() for i in range(5))
    rng = np.random.default_rng(seed + 1)
    m_ = mapping.reshape(8, n, 2)
    copied = np.nonzero(m_[4, :u, 0] == 0)[0]         # e cells copying a
    s1[copied] = e[copied]
    free = np.setdiff1d(np.arange(u), copied)
    e[free] = s1[rng.integers(0, u, free.size)]
    dv = device_fr(H, random_canonical(2 * n, rng))
    t = dv.download()
    dv.free()
    t1, t2 = t[:n].copy(), t[n:].copy()
    t1[u:] = 0
    t2[u:] = 0
    rows = rng.integers(0, u, n)
    f, g = t1[rows].copy(), t2[rows].copy()
    qp = np.zeros((n, 4), dtype=np.uint64)
    qp[:m] = MONT_ONE
    qt = np.zeros((n, 4), dtype=np.uint64)
    qt[1:u:2] = MONT_ONE
    inst = a[:m].copy()
    return (np.concatenate([q, s1, s2, t1, t2, qp, qt]), np.concatenate([a, b, c, d, e, f, g]), [inst],
            mapping)


def bench_prove(args, H):
    """One full proof of the benchmark circuit at k = log_n (default 20):
    5 advice and 3 fixed columns, gate q(ab + c - d), 8-column copy
    permutation, extended domain 4x (2^22).  A step is one
    midnight_hip_plonk_prove call: advice upload, commitments, permutation
    products, numerator and quotient on the 2^22 coset, evaluations and the
    KZG multi-open, returning the proof bytes.  Key and SRS are resident;
    the SRS has 2^max(k, --srs-log) points (--srs-log 20: the serving
    configuration, small circuits over the 2^20-point SRS)."""
    import numerator_graph as G
    k = args.log_n
    srs_log = max(k, args.srs_log)
    rng = np.random.default_rng(args.seed)
    log("building the circuit, SRS (2^%d points) and key for k = %d" % (srs_log, k))
    tau = random_canonical(1, rng)
    tdv = device_fr(H, tau)
    tau_m = tdv.download()[0]
    tdv.free()
    pw = H.DeviceVec.upload(H.fr_shifted_prefix_product(np.tile(tau_m, (1 << srs_log, 1)), MONT_ONE))
    srs = H.MsmContext.from_multiples(pw)
    pw.free()
    gate = ("prod", ("fixed", 0, 0), ("sum", ("sum", ("prod", ("advice", 0, 0), ("advice", 1, 0)),
                                             ("advice", 2, 0)), ("neg", ("advice", 3, 0))))
    graph = G.RawGraphBuffers(G.custom_gates_graph([gate]))
    perm = [(1, i) for i in range(5)] + [(0, i) for i in range(3)]
    inst = None
    if args.arguments:
        A_, F_ = ("advice", 0, 0), ("fixed", 0, 0)
        adv_ = lambda i: ("advice", i, 0)
        fix_ = lambda i: ("fixed", i, 0)
        sub_ = lambda x, y: ("sum", x, ("neg", y))
        gates = [gate, ("prod", fix_(5), sub_(A_, ("instance", 0, 0)))]
        graph = G.RawGraphBuffers(G.custom_gates_graph(gates))
        lk = [([adv_(5), adv_(6)], [fix_(3), fix_(4)]), ([("prod", F_, adv_(4))], [fix_(1)])]
        lookups = [(G.RawGraphBuffers(G.compression_graph(i)), G.RawGraphBuffers(G.compression_graph(t)),
                    G.RawGraphBuffers(G.lookup_graph(i, t))) for i, t in lk]
        r_ = sub_(sub_(adv_(3), ("prod", adv_(0), adv_(1))), adv_(2))
        trash = [(G.RawGraphBuffers(G.trash_graph([r_, ("prod", adv_(4), r_)])), 6)]
        fixed, adv, inst, mapping = synthetic_argument_circuit(H, k, 2000 + k)
        pk = H.PlonkKey(k, srs, 5, fixed, 7, [(c, 0) for c in range(7)], [(c, 0) for c in range(7)], perm,
                        mapping, graph, b"benchmark circuit family with lookups, a trashcan and an instance",
                        instance_cols=1, lookups=lookups, trash=trash)
    else:
        fixed, adv, mapping = synthetic_circuit(H, k, 2000 + k)
        pk = H.PlonkKey(k, srs, 5, fixed, 5, [(c, 0) for c in range(5)], [(c, 0) for c in range(3)], perm,
                        mapping, graph, b"benchmark circuit family")
    for i in range(args.warmup):
        pk.prove(adv, 100 + i, instance=inst)
    H.profile_enable(True)
    lat = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        proof = pk.prove(adv, 200 + i, instance=inst)
        lat.append(time.perf_counter() - t1)
    dt = (time.perf_counter() - t0) / args.steps
    msm_ms, msm_cnt = H.profile_read("msm_total")
    acc_ms, _ = H.profile_read("msm_accumulate")
    stage_ms, _ = H.profile_read("commit_stage")
    num_ms, num_cnt = H.profile_read("numerator")
    H.profile_enable(False)
    peak_measured, _, _ = H.valu_peak()
    peak = PEAK_SPEC_MAC_PER_S
    n = 1 << k
    num_macs = QUOTIENT_FR_MULS_PER_ROW * MACS_PER_FR_MUL * 4 * n
    t_num = num_ms / 1e3 / max(num_cnt, 1)
    lat.sort()
    res = {
        "metric": "k=%d proof latency (proofs/s)" % k, "value": round(1.0 / dt, 3), "unit": "proofs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic satisfying circuit of the benchmark family; SRS [tau^i]G built on the device",
        "config": {"workload": ("full proof, k=%d (2^%d rows), 5 advice / 3 fixed, gate q(ab+c-d), 8-column "
                                "permutation, 4x extended domain" % (k, k)) +
                               (" + advice f g, fixed t1 t2 qp qt, instance p: lookups (f, g) in (t1, t2) and "
                                "q e in s1, trashcan qt, gate qp (a - p)" if args.arguments else ""),
                   "proof_bytes": len(proof),
                   "srs_points": 1 << srs_log},
        "p50_ms": round(lat[len(lat) // 2] * 1e3, 3),
        "stage_ms_per_proof": {"commit_stages_graph": round(stage_ms / args.steps, 3),
                               "msm_total_eager": round(msm_ms / args.steps, 3),
                               "msm_accumulate_eager": round(acc_ms / args.steps, 3),
                               "numerator": round(num_ms / args.steps, 3)},
        "roofline": {"bound": "valu", "kernel": "numerator", "achieved": round(num_macs / t_num / 1e12, 3),
                     "peak": round(peak / 1e12, 3), "peak_measured": round(peak_measured / 1e12, 3),
                     "unit": "TMAC/s", "frac": round(num_macs / t_num / peak, 4), "traffic": None},
    }
    if args.arguments:
        # the numerator's product count above is the benchmark circuit's; the
        # lookup and trash terms add products it does not count
        res["roofline"] = None
    if args.cpu_seconds > 0 and not args.arguments:
        res["cpu_baseline"] = cpu_baseline_prove(args.cpu_seconds)
    pk.free()
    srs.free()
    return res


def cpu_baseline_prove(seconds):
    """The reference-algorithm CPU prover (oracle MSM / FFT / numerator /
    products, protocol restated in tests/plonk_ref.py) on a bounded sample:
    proofs of the same circuit family at k = 14, repeated for about
    `seconds` (the key is built outside the timing)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import plonk_ref
    import synth_circuit
    threads = cpu_threads()
    fc = synth_circuit.FastCircuit(14, seed=14)
    key = plonk_ref.CpuKey(fc, 12345, threads)
    reps, t0 = 0, time.perf_counter()
    while True:
        plonk_ref.cpu_prove(key, fc.advice, reps, threads)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(reps / dt, 4), "unit": "proofs/s", "cores": threads, "kind": "port",
            "sample": "%d proofs at k=14 of the same circuit family (CPU reference prover), %d threads, %.1f s"
                      % (reps, threads, dt)}


SERVE_KS = [12, 14, 16, 18]
SERVE_MIX = [0.5, 0.3, 0.15, 0.05]


def serving_setup(args, H):
    """SRS of 2^20 points [tau^i]G and the four proving keys of the serving
    mix, resident; the one-at-a-time cost of each k (warm)."""
    import numerator_graph as G
    rng = np.random.default_rng(args.seed)
    log("serving: building the 2^20 SRS and the keys for k = 12, 14, 16, 18")
    tau = random_canonical(1, rng)
    tdv = device_fr(H, tau)
    tau_m = tdv.download()[0]
    tdv.free()
    srs_n = 1 << 20
    pw = H.DeviceVec.upload(H.fr_shifted_prefix_product(np.tile(tau_m, (srs_n, 1)), MONT_ONE))
    srs = H.MsmContext.from_multiples(pw)
    pw.free()
    gate = ("prod", ("fixed", 0, 0), ("sum", ("sum", ("prod", ("advice", 0, 0), ("advice", 1, 0)),
                                             ("advice", 2, 0)), ("neg", ("advice", 3, 0))))
    graph = G.RawGraphBuffers(G.custom_gates_graph([gate]))
    perm = [(1, i) for i in range(5)] + [(0, i) for i in range(3)]
    keys, advice, cost = {}, {}, {}
    for k in SERVE_KS:
        fixed, adv, mapping = synthetic_circuit(H, k, 1000 + k)
        keys[k] = H.PlonkKey(k, srs, 5, fixed, 5, [(c, 0) for c in range(5)], [(c, 0) for c in range(3)], perm,
                             mapping, graph, b"benchmark circuit family")
        advice[k] = adv
    for k in SERVE_KS:
        keys[k].prove(advice[k], 1)
        t0 = time.perf_counter()
        for r in range(3):
            keys[k].prove(advice[k], 2 + r)
        cost[k] = (time.perf_counter() - t0) / 3
    mean_cost = sum(cost[k] * p for k, p in zip(SERVE_KS, SERVE_MIX))
    log("serving: single-proof ms %s, one-at-a-time capacity %.1f proofs/s"
        % ({k: round(v * 1e3, 2) for k, v in cost.items()}, 1.0 / mean_cost))
    return {"srs": srs, "keys": keys, "advice": advice, "cost": cost, "mean_cost": mean_cost}


def serve_stream(args, H, st, slots, load, seed):
    """One Poisson stream of args.requests requests (k mix 50/30/15/5 %) at
    `load` x the one-at-a-time capacity, served with `slots` concurrent
    proofs; latency is submission to proof (nearest-rank percentiles)."""
    import threading
    rng = np.random.default_rng(seed)
    mix = rng.choice(SERVE_KS, size=args.requests, p=SERVE_MIX)
    rate = load / st["mean_cost"]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=args.requests))
    keys, advice = st["keys"], st["advice"]
    server = H.ProofServer(slots, args.hbm_budget_gb << 30)
    jobs = [None] * args.requests
    t_start = time.perf_counter()

    def submitter():
        for i in range(args.requests):
            delay = t_start + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            jobs[i] = (time.perf_counter(), server.submit(keys[int(mix[i])], advice[int(mix[i])], 10 + i))

    th = threading.Thread(target=submitter)
    th.start()
    lat = []
    for i in range(args.requests):
        while jobs[i] is None:
            time.sleep(0.0005)
        t_sub, jid = jobs[i]
        server.wait(jid)
        lat.append(time.perf_counter() - t_sub)
    th.join()
    makespan = time.perf_counter() - t_start
    server.free()
    lat = sorted(lat)
    pct = lambda q: lat[min(len(lat) - 1, int(np.ceil(q * len(lat))) - 1)]   # nearest rank
    return {"proofs_per_s": round(args.requests / makespan, 2), "p50_ms": round(pct(0.5) * 1e3, 2),
            "p99_ms": round(pct(0.99) * 1e3, 2), "makespan_s": round(makespan, 3),
            "offered_per_s": round(rate, 2)}


def serving_report(args, H, loads):
    """Batched (args.slots) and one-at-a-time serving of the same Poisson
    streams at each offered load (a multiple of the one-at-a-time
    capacity), and the CPU reference prover on the same mix."""
    st = serving_setup(args, H)
    out = {"slots": args.slots, "requests": args.requests, "hbm_budget_gib": args.hbm_budget_gb,
           "k_mix": {str(k): p for k, p in zip(SERVE_KS, SERVE_MIX)},
           "single_proof_ms": {str(k): round(v * 1e3, 3) for k, v in st["cost"].items()},
           "one_at_a_time_capacity_per_s": round(1.0 / st["mean_cost"], 2), "loads": {}}
    for _ in range(args.warmup):
        serve_stream(args, H, st, args.slots, max(loads), args.seed)
    for load in loads:
        log("serving at %.1f x capacity: one at a time" % load)
        single = serve_stream(args, H, st, 1, load, args.seed + 1)
        log("serving at %.1f x capacity: %d slots" % (load, args.slots))
        batched = serve_stream(args, H, st, args.slots, load, args.seed + 1)
        out["loads"]["%.1f" % load] = {
            "batched": batched, "one_at_a_time": single,
            "speedup": round(batched["proofs_per_s"] / single["proofs_per_s"], 3),
            "p99_ratio": round(batched["p99_ms"] / single["p99_ms"], 3)}
    for k in SERVE_KS:
        st["keys"][k].free()
    st["srs"].free()
    if args.cpu_seconds > 0:
        log("serving: timing the CPU reference prover on the mix")
        out["cpu_baseline"] = cpu_baseline_serving(args.cpu_seconds)
    return out


def cpu_baseline_serving(seconds):
    """The reference-algorithm CPU prover (oracle MSM / FFT / numerator /
    products, protocol in tests/plonk_ref.py) on the serving mix: proofs of
    the same circuit family at k = 12, 14 and 16 (each key built outside
    the timing); the k = 18 cost is the k = 16 cost x 4 (row count).  The
    value is 1 / sum_k p_k t_k, proofs/s one at a time on all threads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import plonk_ref
    import synth_circuit
    threads = cpu_threads()
    t = {}
    budget = {12: 0.2 * seconds, 14: 0.3 * seconds, 16: 0.0}
    for k in (12, 14, 16):
        fc = synth_circuit.FastCircuit(k, seed=k)
        key = plonk_ref.CpuKey(fc, 12345, threads)
        reps, t0 = 0, time.perf_counter()
        while True:
            plonk_ref.cpu_prove(key, fc.advice, reps, threads)
            reps += 1
            if time.perf_counter() - t0 >= budget[k]:
                break
        t[k] = (time.perf_counter() - t0) / reps
    t[18] = 4 * t[16]
    mean = sum(t[k] * p for k, p in zip(SERVE_KS, SERVE_MIX))
    return {"value": round(1.0 / mean, 4), "unit": "proofs/s", "cores": threads, "kind": "port",
            "sample": "CPU prover on the serving mix, %d threads: seconds per proof k=12 %.3f, k=14 %.3f, "
                      "k=16 %.3f measured, k=18 %.3f = 4 x k=16; mix-weighted mean %.3f s"
                      % (threads, t[12], t[14], t[16], t[18], mean)}


def bench_serve(args, H):
    """Serving: a Poisson stream of proof requests for the benchmark circuit
    family at k in {12, 14, 16, 18} (mix 50/30/15/5 %), SRS of 2^20 points
    and the four proving keys resident, HBM-budgeted admission.  The same
    stream is served one proof at a time (one slot) and with continuous
    batching (--slots); the value is the batched proofs/s at --load, with
    p50/p99 latency (submission to proof) for both, and the same at 0.9 x
    the one-at-a-time capacity."""
    rep = serving_report(args, H, sorted({args.load, 0.9}, reverse=True))
    main = rep["loads"]["%.1f" % args.load]
    res = {
        "metric": "serving proofs/s (k in {12,14,16,18}, Poisson arrivals), batched vs one-at-a-time",
        "value": main["batched"]["proofs_per_s"], "unit": "proofs/s", "n_gpus": 1, "steps": args.requests,
        "warmup": args.warmup, "ms_per_step": round(1e3 / main["batched"]["proofs_per_s"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic satisfying circuits of the benchmark family, Poisson arrivals from the seed",
        "config": {"workload": "proof server, %d requests, k mix 12/14/16/18 at 50/30/15/5%%, SRS 2^20 resident, "
                               "%d slots, HBM budget %d GiB, offered load %.1f x one-at-a-time capacity"
                               % (args.requests, args.slots, args.hbm_budget_gb, args.load)},
        "batched": main["batched"], "one_at_a_time": main["one_at_a_time"], "speedup": main["speedup"],
        "serving": rep,
    }
    if "cpu_baseline" in rep:
        res["cpu_baseline"] = rep["cpu_baseline"]
    return res


QUOTIENT_FR_MULS_PER_ROW = 52      # gate 3, permutation 49 (three sets over 8 columns, chunk 3)
MACS_PER_FR_MUL = 136              # 2 * 8^2 + 8 32x32->64 products per Montgomery product
QUOTIENT_BYTES_PER_ROW = 23 * 32   # 5 advice + 3 fixed + 8 sigma + 3 z + l0, l_last, l_active + output


def bench_quotient(args, H):
    """Quotient numerator of the benchmark circuit (k = 20: 5 advice,
    3 fixed, gate q*(a*b + c - d), 8 permutation columns in 3 sets) on its
    extended domain, 2^log_n rows (default 22).  A step is one resident
    numerator_context_eval: inputs resident in HBM, output a new device
    vector.  Column values are uniform synthetic Fr (the kernel's work does
    not depend on them)."""
    import numerator_graph as G
    log_n = args.log_n if args.log_n != 20 else 22
    n = 1 << log_n
    rng = np.random.default_rng(args.seed)
    log("generating %d x 2^%d synthetic columns" % (8 + 14, log_n))
    cols = device_fr(H, random_canonical(n * 8, rng))          # advice a..e | z_0..z_2
    key = [random_canonical(n, rng) for _ in range(14)]         # q s1 s2 | sigma x8 | l0 l_last l_active
    scal = random_canonical(8, rng)
    gate = ("prod", ("fixed", 0, 0), ("sum", ("sum", ("prod", ("advice", 0, 0), ("advice", 1, 0)),
                                              ("advice", 2, 0)), ("neg", ("advice", 3, 0))))
    graph = G.RawGraphBuffers(G.custom_gates_graph([gate]))
    perm_cols = [(1, i) for i in range(5)] + [(0, i) for i in range(3)]
    ctx = H.NumeratorContext(
        n, n >> 2, fixed=np.concatenate(key[:3]), fixed_lagrange=np.concatenate(key[:3])[: 3 * (n >> 2)],
        l0=key[11], l_last=key[12], l_active=key[13], permutation_columns=perm_cols,
        permutation_pk_cosets=np.concatenate(key[3:11]),
        permutation_pk_lagrange=np.concatenate(key[3:11])[: 8 * (n >> 2)], custom_graph=graph,
        omega=scal[0], trace_omega=scal[1], g_coset=scal[2], delta_const=scal[3], rot_scale=4,
        last_rotation=-6, chunk_len=3)
    del key
    ch = dict(beta=scal[4], gamma=scal[5], theta=scal[6], trash_challenge=scal[7], y=scal[0])

    def step():
        out = ctx.eval(cols, 8, advice_offset=0, advice_cols=5, permutation_offset=5, permutation_sets=3, **ch)
        out.free()

    for _ in range(args.warmup):
        step()
    H.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = (time.perf_counter() - t0) / args.steps
    k_ms, k_cnt = H.profile_read("numerator")
    H.profile_enable(False)
    t_kernel = k_ms / 1e3 / max(k_cnt, 1)
    peak_measured, _, _ = H.valu_peak()
    peak = PEAK_SPEC_MAC_PER_S
    macs = QUOTIENT_FR_MULS_PER_ROW * MACS_PER_FR_MUL * n
    bytes_alg = QUOTIENT_BYTES_PER_ROW * n
    res = {
        "metric": "quotient numerator Mrows/s @2^%d extended rows" % log_n,
        "value": round(n / dt / 1e6, 3), "unit": "Mrows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic uniform Fr columns, resident in HBM",
        "config": {"workload": "numerator of the k=%d circuit (5 advice, 3 fixed, gate q(ab+c-d), "
                               "8-column permutation in 3 sets) on 2^%d extended rows" % (log_n - 2, log_n),
                   "rows": n},
        "roofline": {"bound": "valu", "kernel": "numerator",
                     "achieved": round(macs / t_kernel / 1e12, 3), "peak": round(peak / 1e12, 3),
                     "peak_measured": round(peak_measured / 1e12, 3),
                     "unit": "TMAC/s", "frac": round(macs / t_kernel / peak, 4), "traffic": None},
        "hbm": {"achieved": round(bytes_alg / t_kernel / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(bytes_alg / t_kernel / HBM_PEAK, 4)},
        "kernel_ms_avg": round(t_kernel * 1e3, 3),
    }
    if args.cpu_seconds > 0:
        log("timing the CPU baseline")
        res["cpu_baseline"] = cpu_baseline_quotient(args.cpu_seconds)
    ctx.free()
    cols.free()
    return res


def cpu_baseline_quotient(seconds):
    """Oracle evaluate_numerator (row-parallel C restatement) on a bounded
    sample: the same circuit shape over 2^18 rows, repeated for about
    `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    import synth_circuit
    threads = cpu_threads()
    log_s = 18
    inp = synth_circuit.random_inputs(1 << log_s, 5)
    oracle.numerator(**dict(inp, n=1024, fixed=inp["fixed"][:3 * 1024], advice=inp["advice"][:5 * 1024],
                            l0=inp["l0"][:1024], l_last=inp["l_last"][:1024], l_active=inp["l_active"][:1024],
                            permutation_products=inp["permutation_products"][:3 * 1024],
                            permutation_pk_cosets=inp["permutation_pk_cosets"][:8 * 1024]), threads=threads)
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.numerator(**inp, threads=threads)
        reps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round((1 << log_s) * reps / dt / 1e6, 4), "unit": "Mrows/s", "cores": threads, "kind": "port",
            "sample": "%d x numerator over 2^%d rows of the same circuit shape, oracle evaluate_numerator, "
                      "%d threads, %.1f s" % (reps, log_s, threads, dt)}


POSEIDON_ROUNDS = 68
POSEIDON_STEPS_PER_ROUND = 3      # dependent Fr products per round on the chain kernel's critical path


def bench_poseidon(args, H):
    """Poseidon hash chain (width 3, 8 full + 60 partial rounds, x^5) of
    --chain-len dependent hashes through fr_poseidon_hash_chain_witness, the
    form the circuit-witness preprocessing calls: a step is one call --
    parameters uploaded, the chain, its 256 witness values per hash copied to
    the host.  MDS and round constants are uniform synthetic Fr (the work
    does not depend on them).  The chain is sequential, so the bound is the
    latency of one wavefront's dependent Fr products: `roofline` compares the
    critical-path products per second (3 per round, 136 rounds per hash)
    with the single-wave dependent-product rate of the radix-2^29 product
    the kernel uses, measured on the same device."""
    length = args.chain_len
    rng = np.random.default_rng(args.seed)
    # values in [0, r) are the Montgomery forms of uniform field elements
    mds_m, rc_m, init_m = random_canonical(9, rng), random_canonical(204, rng), random_canonical(4, rng)

    def step():
        return H.fr_poseidon_hash_chain_witness(length, init_m[0], init_m[1], init_m[2], init_m[3], mds_m, rc_m)

    for _ in range(args.warmup):
        step()
    H.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = (time.perf_counter() - t0) / args.steps
    k_ms, k_cnt = H.profile_read("poseidon_chain")
    H.profile_enable(False)
    t_kernel = k_ms / 1e3 / max(k_cnt, 1)
    lat = H.fr29_product_latency()              # seconds per dependent (radix-2^29) product on one wave
    products = POSEIDON_STEPS_PER_ROUND * POSEIDON_ROUNDS * 2 * length
    res = {
        "metric": "Poseidon hash-chain hashes/s (chain of %d, with witness)" % length,
        "value": round(length / dt, 1), "unit": "hashes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic uniform Fr parameters and chain inputs",
        "config": {"workload": "Poseidon chain of %d hashes (2 permutations each), witness to host" % length,
                   "hashes": length},
        "roofline": {"bound": "valu-latency", "kernel": "k_poseidon_chain",
                     "achieved": round(products / t_kernel / 1e6, 3), "peak": round(1.0 / lat / 1e6, 3),
                     "unit": "M dependent Fr products/s", "frac": round(products * lat / t_kernel, 4),
                     "traffic": None},
        "kernel_ms_avg": round(t_kernel * 1e3, 3),
        "us_per_hash": round(t_kernel / length * 1e6, 2),
        "fr_product_latency_ns": round(lat * 1e9, 1),
    }
    if args.cpu_seconds > 0:
        log("timing the CPU baseline")
        res["cpu_baseline"] = cpu_baseline_poseidon(args.cpu_seconds, mds_m, rc_m, init_m)
    return res


def cpu_baseline_poseidon(seconds, mds_m, rc_m, init_m):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    sample = 256
    reps, t0 = 0, time.perf_counter()
    while True:
        O.poseidon_hash_chain(sample, init_m[0], init_m[1], init_m[2], init_m[3], mds_m, rc_m, witness=True)
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    return {"value": round(reps * sample / dt, 1), "unit": "hashes/s", "cores": 1, "kind": "port",
            "sample": "%d chains of %d hashes with witness, oracle poseidon_hash_chain, 1 thread, %.1f s"
                      % (reps, sample, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="msm", choices=["msm", "ntt", "batch", "quotient", "kzg", "serve", "prove",
                                                         "poseidon"])
    ap.add_argument("--log-n", type=int, default=20)
    ap.add_argument("--srs-log", type=int, default=0, help="prove: SRS of 2^max(log-n, srs-log) points")
    ap.add_argument("--msms-per-step", type=int, default=4)
    ap.add_argument("--skewed", action="store_true")
    ap.add_argument("--arguments", action="store_true",
                    help="prove: the circuit with two lookups, a trashcan and an instance column")
    ap.add_argument("--chain-len", type=int, default=1024, help="poseidon: hashes in the chain")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--requests", type=int, default=400, help="serving: requests in the stream")
    ap.add_argument("--slots", type=int, default=12, help="serving: concurrent proofs")
    ap.add_argument("--load", type=float, default=4.0, help="serving: offered load / one-at-a-time capacity")
    ap.add_argument("--hbm-budget-gb", type=int, default=64, help="serving: HBM admission budget")
    ap.add_argument("--pmc-copy-mib", type=int, default=0,
                    help="msm: run one device copy of this many MiB before timing (counter calibration)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU-baseline sample length (0 disables)")
    ap.add_argument("--no-serving", dest="serving", action="store_false",
                    help="msm: skip the serving section of the default run")
    args = ap.parse_args()
    if args.workload != "msm" or (args.log_n, args.msms_per_step, args.skewed) != (20, 4, False):
        args.serving = False
    if args.gpus != 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        return
    if args.workload == "serve" or args.serving:
        # one hardware queue per lane stream (the proof server's slots plus
        # the caller's lane 0); HIP's default of 4 would make lanes share
        # queues and serialise their kernels.  Read when HIP starts, so it is
        # set before the library is loaded.
        want = 16 if args.slots + 1 <= 16 else 32
        os.environ["GPU_MAX_HW_QUEUES"] = str(max(int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), want))
    import midnight_hip as H
    H.lib()
    run = {"msm": bench_msm, "ntt": bench_ntt, "batch": bench_batch, "quotient": bench_quotient, "kzg": bench_kzg, "serve": bench_serve, "prove": bench_prove,
           "poseidon": bench_poseidon}[args.workload]
    res = run(args, H)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
