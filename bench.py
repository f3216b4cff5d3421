#!/usr/bin/env python3
"""DSOFT + iDSOFT throughput on one MI355X.

Workload (default): the plain FP64 single-signal round trip at bandwidth
B = 128 -- seeded coefficients f^l_{M,M'} (0 <= l < 128, uniform[-1,1]/(l+1),
2,796,160 complex128 values) -> iDSOFT -> 256^3 complex128 grid (268 MB) ->
DSOFT -> coefficients.  One step = one iDSOFT + one DSOFT on device-resident
buffers; value = round-trip pairs per second.

Also reported: the roofline figure of the dominant kernel, whose average
duration comes from HIP events recorded in its launches' dispatch packets on
the launch stream over the timed region (the only kernel timed there); a
per-kernel breakdown from a separate pass with every kernel timed; per-step
latency percentiles, the round-trip error, and the CPU oracle (C/OpenMP
restatement of the reference) timed on this host on a bounded sample of the
same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--bandwidth B]
"""
import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
FP32_MFMA_PEAK_TFS = 157.3  # MI355X dense FP32 matrix, published
FP64_PEAK_TFS = 78.6  # MI355X dense FP64 matrix and FP64 vector, published


def algorithmic_bytes_per_step(kernel, B, e=16):
    """Algorithmic HBM bytes one transform moves through this kernel (summed
    over its launches).  N = 2B, e = bytes per complex value.
    FFT stage per direction: 2 N^3 e (one read + one write of the grid),
    split evenly between its row and column passes.
    DWT stage per direction: (N-1)^2 N e pencil bytes + B(4B^2-1)/3 e
    coefficient bytes; up to N = 256 the inverse DWT launch also copies the
    (2N-1) N stale Nyquist entries (one read and one write each)."""
    N = 2 * B
    if kernel.startswith("fft_"):
        return N ** 3 * e
    if kernel.startswith("dwt_"):
        nyq = 2 * (2 * N - 1) * N * e if kernel == "dwt_inv" and N <= 256 else 0
        return (N - 1) ** 2 * N * e + B * (4 * B * B - 1) // 3 * e + nyq
    if kernel == "nyquist_gather":
        return 2 * (2 * N - 1) * N * e
    return 0


def algorithmic_flops_per_step(kernel, B):
    """Algorithmic FP64 flops of one transform in this kernel.  A DWT direction
    contracts every one of the B(4B^2-1)/3 coefficients with its N-node
    Wigner-d vector (complex x real: 4 flop per node), and builds those vectors
    per transform as the reference does (fn_dsoft.cpp:227, fn_idsoft.cpp:224)
    with the three-term recurrence: 5 flop per node, degree and canonical pair
    (sum over pairs M >= M' of the B - M degrees).  The FFT passes' flops are
    left out: at 5 log2 N flop per 16-byte point they are HBM-bound by two
    orders of magnitude."""
    N = 2 * B
    if kernel.startswith("dwt_"):
        rows = sum((M + 1) * (B - M) for M in range(B))
        return 4 * N * (B * (4 * B * B - 1) // 3) + 5 * N * rows
    return 0


def roofline_of(alg_bytes, alg_flops, us):
    """The binding roofline of one launch: the larger of its HBM floor
    (algorithmic bytes at 8 TB/s) and its FP64 floor (algorithmic flops at
    78.6 TFLOP/s, the matrix and the vector figure alike), with the achieved
    rate and fraction against that peak, and both floors."""
    t_hbm = alg_bytes / (HBM_PEAK_GBS * 1e9) * 1e6
    t_fp64 = alg_flops / (FP64_PEAK_TFS * 1e12) * 1e6
    hbm = {"achieved": round(alg_bytes / (us * 1e-6) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(alg_bytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4), "floor_us": round(t_hbm, 2)}
    if not alg_flops or t_fp64 <= t_hbm:
        return "hbm", hbm, None
    fp64 = {"achieved": round(alg_flops / (us * 1e-6) / 1e12, 2), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
            "frac": round(alg_flops / (us * 1e-6) / 1e12 / FP64_PEAK_TFS, 4), "floor_us": round(t_fp64, 2)}
    return "mfma", fp64, hbm


def measured_traffic(B, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/pmc_traffic_f64_B<B>.json), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic_f64_B%d.json" % B)
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def measured_pipes(B):
    """Per-kernel SQ instruction counters of this workload from the committed
    rocprofv3 summary (profiles/pmc_pipes_f64_B<B>.json), or {}."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_pipes_f64_B%d.json" % B)
    try:
        with open(path) as f:
            return json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return {}


def batched_counters(B, nsig):
    """Committed rocprofv3 counters of the batched kernels at this workload
    (profiles/pmc_batched.json: 64 signals at B=128, 512 at B=64), or {}."""
    key = {(128, 64): "64xB128", (64, 512): "512xB64"}.get((B, nsig))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_batched.json")
    if not key:
        return {}
    try:
        with open(path) as f:
            return json.load(f)["workloads"][key]["kernels"]
    except (OSError, KeyError, ValueError):
        return {}


def pipe_report(B, kernels):
    """For each kernel with committed SQ counters: the SIMD issue floor (VALU
    and MFMA cycles the counters give, spread over every SIMD at the peak
    clock) against the live average launch time, and the FP64 rate the
    counted MFMA and FMA work reaches at that time."""
    res = {}
    for name, c in measured_pipes(B).items():
        if name not in kernels:
            continue
        us = kernels[name]["avg_launch_us"]
        flop = c["mfma_f64_flop"] + c["valu_f64_fma_flop"]
        res[name] = {"issue_floor_us": c["issue_floor_us"], "avg_launch_us": us,
                     "issue_frac": round(c["issue_floor_us"] / us, 4),
                     "fp64_TFLOPs": round(flop / (us * 1e-6) / 1e12, 2),
                     "fp64_frac": round(flop / (us * 1e-6) / 1e12 / FP64_PEAK_TFS, 4)}
    return res


def cpu_model():
    return cpu_info().get("model name", "unknown")


def cpu_info():
    """The first processor's /proc/cpuinfo fields that name the host CPU, plus
    the count of logical processors and the threads this process may use."""
    keep = ("vendor_id", "cpu family", "model", "model name", "stepping", "microcode", "cpu MHz", "cache size",
            "siblings", "cpu cores", "address sizes")
    out, n = {}, 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                if k == "processor":
                    n += 1
                elif n == 1 and k in keep:
                    out[k] = v
    except OSError:
        pass
    out["logical_processors"] = n
    try:
        out["usable_threads"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return out


def cpu_baseline(B, pairs, seed, sweep_B=64):
    """The oracle (C/OpenMP restatement of the reference's DSOFT/IDSOFT) on this
    host, following the reference's benchmark protocol
    (benchmark/benchmark_dsoft_for.cpp:44-48,135-223): warm the OpenMP pool,
    build the input once, time forward and inverse separately over `pairs`
    round trips with all host threads (avg/min/max), then one round trip with
    one thread for the serial time, speedup and parallel efficiency; then the
    reference's thread sweep at bandwidth sweep_B (thread_sweep)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import numpy as np
    import oracle_bind as O
    threads = O.lib().oracle_max_threads()
    c = O.seeded_coeffs(B, seed)
    O.dsoft(4, O.idsoft(4, O.seeded_coeffs(4, 1)))  # warm the OpenMP pool

    def run(n, th):
        t_inv, t_fwd, r = [], [], None
        for _ in range(n):
            t0 = time.perf_counter()
            g = O.idsoft(B, c, threads=th)
            t1 = time.perf_counter()
            r = O.dsoft(B, g, threads=th)
            t2 = time.perf_counter()
            t_inv.append(t1 - t0)
            t_fwd.append(t2 - t1)
        return t_inv, t_fwd, r

    t_inv, t_fwd, r = run(pairs, threads)
    err = float(np.linalg.norm(r - c) / np.linalg.norm(c))
    progress("CPU baseline: %d round trips on %d threads done" % (pairs, threads))
    s_inv, s_fwd, _ = run(1, 1)
    progress("CPU baseline: serial round trip done")
    par = (sum(t_inv) + sum(t_fwd)) / pairs
    ser = s_inv[0] + s_fwd[0]
    sweep = thread_sweep(O, sweep_B, threads, seed) if sweep_B else None
    progress("CPU baseline: thread sweep done")
    fft = fft_share(O, np, B, O.idsoft(B, c, threads=threads), threads, par)
    return {
        "value": 1.0 / par,
        "unit": "DSOFT+iDSOFT pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d round trips at B=%d with %d threads plus 1 with 1 thread (same seeded input); oracle "
                  "C/OpenMP restatement of the reference (long double DWT, own long double FFT in place of FFTW)"
                  % (pairs, B, threads),
        "cpu_model": cpu_model(),
        "forward_s": {"avg": sum(t_fwd) / pairs, "min": min(t_fwd), "max": max(t_fwd)},
        "inverse_s": {"avg": sum(t_inv) / pairs, "min": min(t_inv), "max": max(t_inv)},
        "serial_pair_s": ser,
        "speedup": ser / par,
        "efficiency": ser / par / threads,
        "roundtrip_rel_l2": err,
        "cpu_info": cpu_info(),
        "thread_sweep": sweep,
        "fft_share": fft,
    }


def fft_share(O, np, B, grid, threads, pair_s, reps=2):
    """The FFT stage of the CPU baseline against an FFTW-class FFT.  The
    reference runs its slice FFTs through FFTW (src/fftw_wrapper.cpp:47-172);
    FFTW is not installed here, so the oracle uses its own long-double FFT.
    This times that FFT stage (one 2D FFT of every slice, as each direction
    runs once) against scipy.fft's pocketfft in double precision on the same
    threads, and estimates the round trip with the latter in its place."""
    N = 2 * B
    a = np.ascontiguousarray(grid.reshape(N, N, N))
    t0 = time.perf_counter()
    for _ in range(reps):
        O.fft2_layers(a, -1, threads=threads)
    t_oracle = (time.perf_counter() - t0) / reps
    try:
        import scipy.fft as sf
    except ImportError:
        return {"oracle_fft_stage_s": t_oracle, "note": "scipy not importable: no FFTW-class comparison"}
    sf.fft2(a, axes=(1, 2), workers=threads)
    t0 = time.perf_counter()
    for _ in range(reps):
        sf.fft2(a, axes=(1, 2), workers=threads)
    t_pocket = (time.perf_counter() - t0) / reps
    est = pair_s - 2 * t_oracle + 2 * t_pocket
    return {"oracle_fft_stage_s": t_oracle, "pocketfft_stage_s": t_pocket, "threads": threads,
            "pair_s": pair_s, "pair_s_with_pocketfft": est, "pairs_per_s_with_pocketfft": 1.0 / est,
            "note": "one 2D FFT of every slice per direction; the estimate replaces the oracle's two FFT stages "
                    "per round trip by scipy.fft's (pocketfft, double precision, an FFTW-class FFT)"}


def thread_sweep(O, B, max_threads, seed, runs=2):
    """The reference's speedup and efficiency drivers
    (benchmark/benchmark_dsoft_for_speedup.cpp:51-160,
    benchmark_dsoft_for_efficiency.cpp:160): forward and inverse at every
    thread count 1..max_threads (the serial run first), `runs` round trips
    each, speedup = serial / t and efficiency = serial / (threads * t) per
    direction.  Run at a bandwidth small enough that the whole sweep stays a
    few seconds of host time."""
    c = O.seeded_coeffs(B, seed)
    g = O.idsoft(B, c)
    rows = []
    for th in range(1, max_threads + 1):
        tf = ti = 0.0
        for _ in range(runs):
            t0 = time.perf_counter()
            O.idsoft(B, c, threads=th)
            t1 = time.perf_counter()
            O.dsoft(B, g, threads=th)
            t2 = time.perf_counter()
            ti += t1 - t0
            tf += t2 - t1
        rows.append((th, tf / runs, ti / runs))
    sf, si = rows[0][1], rows[0][2]
    return {
        "bandwidth": B, "runs_per_count": runs,
        "threads": [r[0] for r in rows],
        "forward_s": [round(r[1], 6) for r in rows], "inverse_s": [round(r[2], 6) for r in rows],
        "forward_speedup": [round(sf / r[1], 3) for r in rows], "inverse_speedup": [round(si / r[2], 3) for r in rows],
        "forward_efficiency": [round(sf / r[1] / r[0], 3) for r in rows],
        "inverse_efficiency": [round(si / r[2] / r[0], 3) for r in rows],
    }


def cpu_other_configs(seed):
    """The oracle on this host at the other CPU-runnable workloads, all host
    threads, bounded samples: the B=32 round trip (five round trips), and the
    512-signal B=64 batch as sequential complex128 round trips (four signals
    of the 512; the CPU path has no complex64 or batched form)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import numpy as np
    import oracle_bind as O
    threads = O.lib().oracle_max_threads()
    res = {}
    for name, B, n in (("b32_roundtrip", 32, 5), ("b64_batch_sample", 64, 4)):
        t_inv = t_fwd = 0.0
        err = 0.0
        for i in range(n):
            c = O.seeded_coeffs(B, seed + i)
            t0 = time.perf_counter()
            g = O.idsoft(B, c, threads=threads)
            t1 = time.perf_counter()
            r = O.dsoft(B, g, threads=threads)
            t2 = time.perf_counter()
            t_inv += t1 - t0
            t_fwd += t2 - t1
            err = max(err, float(np.linalg.norm(r - c) / np.linalg.norm(c)))
        res[name] = {"bandwidth": B, "pairs": n, "pairs_per_s": n / (t_inv + t_fwd), "cores": threads,
                     "forward_s_avg": t_fwd / n, "inverse_s_avg": t_inv / n, "roundtrip_rel_l2_max": err}
    res["note"] = ("oracle (C/OpenMP restatement of the reference), complex128; b64_batch_sample: %d of the 512 "
                   "signals of the batched B=64 workload, one after another" % res["b64_batch_sample"]["pairs"])
    return res


def batch_dwt_flops(B):
    """Useful FP32 flops of one signal's batched DWT in one direction: every
    active (pair, case) contracts R = B - M degrees against N = 2B betas for
    the real and the imaginary part."""
    N = 2 * B
    f = 0
    for M in range(B):
        for Mp in range(M + 1):
            nc = 1 if M == 0 else (4 if (Mp == 0 or Mp == M) else 8)
            f += nc * 2 * (B - M) * N * 2
    return f


def batched_leg(torch, P, ctx, B, nsig, steps, warmup, seed, check=0, prec=0):
    """The batched mixed-precision round trip over nsig complex64 signals,
    with the Wigner contraction in FP32 MFMA (prec 0) or in split-bf16 x3 / x6
    products on bf16 MFMA (prec 3 / 6, context tune key batch_precision).

    Every signal's result is checked against the plain FP64 GPU path (which
    the parity tests pin to the CPU oracle at <= 1e-11) on the same input:
    inverse grids against FP64 iDSOFT of the complex64 coefficients, forward
    coefficients against FP64 DSOFT of the complex64 grid."""
    N = 2 * B
    n = P.coeff_count(B)
    stream = torch.cuda.current_stream()
    c32 = P.seeded_coeffs_torch(B, [seed + i for i in range(nsig)], "cuda", dtype=torch.complex64)
    grids = torch.empty((nsig, N ** 3), dtype=torch.complex64, device="cuda")
    back = torch.empty_like(c32)
    ctx.warm(B)
    ctx.tune("batch_precision", prec)

    def step():
        ctx.idsoft_batch_f32(B, nsig, c32, grids, stream=stream)
        ctx.dsoft_batch_f32(B, nsig, grids, back, stream=stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ctx.timing(True)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.timing(False)
    ctx.tune("batch_precision", 0)
    lat = sorted(ev[i][0].elapsed_time(ev[i][1]) for i in range(steps))

    kernels = {}
    dwt_flops = batch_dwt_flops(B) * nsig
    for name, (cnt, tot_ms) in ctx.timing_collect().items():
        k = {"launches_per_step": cnt // steps, "us_per_step": round(tot_ms / steps * 1e3, 2),
             "avg_launch_us": round(tot_ms / cnt * 1e3, 2)}
        if name.startswith("batch_dwt_"):
            tf = dwt_flops / (tot_ms / steps / 1e3) / 1e12
            k["useful_TFLOPs"] = round(tf, 2)
            k["frac_fp32_mfma"] = round(tf / FP32_MFMA_PEAK_TFS, 4)
        if name.startswith("batch_fft_"):
            # a pass reads and writes every signal's grid once: its own traffic
            gbs = nsig * N ** 3 * 8 * 2 / (tot_ms / steps / 1e3) / 1e9
            k["pass_traffic_GBps"] = round(gbs, 1)
            k["pass_traffic_frac_hbm"] = round(gbs / HBM_PEAK_GBS, 4)
        kernels[name] = k
    # FFT stage per direction, algorithmic: one read and one write of the grids
    # (2 N^3 e per signal) over the time of both passes of that direction
    # committed counters of these kernels (traffic per launch from FETCH_SIZE /
    # WRITE_SIZE, MFMA busy cycles), set against this run's launch times: the
    # MFMA pipes' busy fraction of every SIMD at the 2.4 GHz peak clock over
    # the live launch, and the leg's counter traffic against its algorithmic
    # bytes per round trip
    cnt = batched_counters(B, nsig)
    counter_summary = None
    if cnt:
        traffic_step = 0
        for name, k in kernels.items():
            key = name
            if name.startswith("batch_fft_fwd"):
                key = "batch_fft_fwd_rows+cols"
            elif name.startswith("batch_fft_inv"):
                key = "batch_fft_inv_cols+rows"
            c = cnt.get(key)
            if name == "batch_coeff_transpose":
                # one read + one write of the coefficient blocks per launch (counters: 1.000x)
                traffic_step += 2 * 2 * nsig * (B * (4 * B * B - 1) // 3) * 8
            if not c:
                continue
            k["counter_traffic_bytes_per_launch"] = int(c["traffic_bytes"])
            k["counter_traffic_over_alg"] = c["traffic_over_alg"]
            traffic_step += c["traffic_bytes"] * k["launches_per_step"]
            if "mfma_busy_cycles" in c:
                k["mfma_busy_counter_run"] = c["mfma_busy_run"]
                k["mfma_busy_live"] = round(c["mfma_busy_cycles"] / (1024 * 2.4e9 * k["avg_launch_us"] * 1e-6), 4)
        alg_step = nsig * 2 * (2 * N ** 3 * 8 + ((N - 1) ** 2 * N + B * (4 * B * B - 1) // 3) * 8)
        counter_summary = {"traffic_bytes_per_roundtrip": int(traffic_step), "alg_bytes_per_roundtrip": alg_step,
                           "traffic_over_alg": round(traffic_step / alg_step, 3),
                           "note": "counter figures from profiles/pmc_batched.json (rocprofv3 2 x FETCH_SIZE + "
                                   "WRITE_SIZE per launch, SQ_VALU_MFMA_BUSY_CYCLES); mfma_busy_live = committed "
                                   "busy cycles / (1024 SIMDs x 2.4 GHz x this run's launch time)"}
    # (one launch of the one-pass slice kernel at 2B = 64 and 128, else two passes)
    fft_stage = {}
    for d, alts in (("forward", (("batch_fft_fwd_slice",), ("batch_fft_fwd_rows", "batch_fft_fwd_cols"))),
                    ("inverse", (("batch_fft_inv_slice",), ("batch_fft_inv_cols", "batch_fft_inv_rows")))):
        for names in alts:
            if all(x in kernels for x in names):
                us = sum(kernels[x]["us_per_step"] for x in names)
                gbs = nsig * 2 * N ** 3 * 8 / (us * 1e-6) / 1e9
                fft_stage[d] = {"us": round(us, 2), "alg_GBps": round(gbs, 1),
                                "alg_frac_hbm": round(gbs / HBM_PEAK_GBS, 4), "kernels": list(names)}
                break

    # error of every signal against the FP64 path on the same input
    g64 = torch.empty(N ** 3, dtype=torch.complex128, device="cuda")
    c64 = torch.empty(n, dtype=torch.complex128, device="cuda")
    e_inv, e_fwd, e_rt = [], [], []
    nrm = torch.linalg.vector_norm
    nchk = nsig if check <= 0 else min(check, nsig)
    for i in range(nchk):
        g64.zero_()  # iDSOFT reads the Nyquist entries of the grid it is given
        ctx.idsoft_dev(B, c32[i].to(torch.complex128), g64, stream=stream)
        e_inv.append((nrm(grids[i].to(torch.complex128) - g64) / nrm(g64)).item())
        ctx.dsoft_dev(B, grids[i].to(torch.complex128), c64, stream=stream)
        e_fwd.append((nrm(back[i].to(torch.complex128) - c64) / nrm(c64)).item())
        ref = c32[i].to(torch.complex128)
        e_rt.append((nrm(back[i].to(torch.complex128) - ref) / nrm(ref)).item())
    elapsed = t1 - t0
    contraction = {0: "fp32 MFMA", 3: "split-bf16 x3 on bf16 MFMA (fp32 accumulation)",
                   6: "split-bf16 x6 on bf16 MFMA (fp32 accumulation)"}[prec]
    return {
        "bandwidth": B, "signals": nsig,
        "precision": "fp32 FFT + %s DWT, fp64 tables/weights/twiddles" % contraction,
        "pairs_per_s": nsig * steps / elapsed, "transforms_per_s": 2 * nsig * steps / elapsed,
        "ms_per_batch_roundtrip": elapsed / steps * 1e3, "steps": steps, "warmup": warmup,
        "latency_ms": {"p50": lat[len(lat) // 2], "p99": lat[min(len(lat) - 1, int(0.99 * len(lat)))]},
        "rel_l2_vs_fp64": {"inverse_max": max(e_inv), "forward_max": max(e_fwd), "roundtrip_max": max(e_rt),
                           "inverse_median": sorted(e_inv)[nchk // 2], "forward_median": sorted(e_fwd)[nchk // 2],
                           "signals_checked": nchk, "bound": 1e-6},
        "kernels": kernels,
        "fft_stage": fft_stage,
        "counters": counter_summary,
    }


SERVE_BANDWIDTHS = [16, 32, 64, 128]
SERVE_PROBS = [0.4, 0.3, 0.2, 0.1]


def serving_leg(torch, P, ctx, seconds, seed, budget_bytes, slots, max_batch, util=0.7, rate=None):
    """Seeded Poisson request stream through the request server.

    Bandwidths {16, 32, 64, 128} with probabilities {0.4, 0.3, 0.2, 0.1};
    50/50 forward DSOFT on an iDSOFT-made complex64 grid and inverse iDSOFT on
    seeded coefficients; inputs resident in HBM, outputs into rings of device
    buffers.  The offered rate is util x the capacity measured by a saturating
    burst of the same mix (unless given).  Reports p50/p99 latency (submit to
    completion, host clock), completed transforms/s and the error of every
    ring slot's last result against the plain FP64 path on the same input."""
    import numpy as np
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(seed)
    nin = 4
    ring = {16: 256, 32: 256, 64: 96, 128: 32}
    inputs, outs, last = {}, {}, {}
    for B in SERVE_BANDWIDTHS:
        N, n = 2 * B, P.coeff_count(B)
        c = P.seeded_coeffs_torch(B, [seed + 100 * B + i for i in range(nin)], "cuda", dtype=torch.complex64)
        g = torch.empty((nin, N ** 3), dtype=torch.complex64, device="cuda")
        ctx.idsoft_batch_f32(B, nin, c, g, stream=stream)
        inputs[(B, P.PFS_REQ_INVERSE)] = c
        inputs[(B, P.PFS_REQ_FORWARD)] = g
        outs[(B, P.PFS_REQ_INVERSE)] = torch.empty((ring[B], N ** 3), dtype=torch.complex64, device="cuda")
        outs[(B, P.PFS_REQ_FORWARD)] = torch.empty((ring[B], n), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    srv = P.Server(ctx, SERVE_BANDWIDTHS, slots=slots, max_batch=max_batch, hbm_budget=budget_bytes)
    pos = {key: 0 for key in outs}
    busy = {}  # (key, ring slot) -> ticket not yet collected
    owner = {}  # ticket -> (key, ring slot)
    lat = []

    def collect():
        for t, l, st in srv.collect():
            if st != 0:
                raise RuntimeError("serving request failed with status %d" % st)
            lat.append(l)
            busy.pop(owner.pop(t), None)

    def submit(i):
        B = int(rng.choice(SERVE_BANDWIDTHS, p=SERVE_PROBS))
        kind = int(rng.integers(0, 2))
        key = (B, kind)
        r = pos[key]
        pos[key] = (r + 1) % ring[B]
        while (key, r) in busy:  # ring slot still in flight
            collect()
        j = i % nin
        t = srv.submit(kind, B, inputs[key][j], outs[key][r])
        busy[(key, r)] = t
        owner[t] = (key, r)
        last[(key, r)] = j

    # capacity: a saturating burst of the same mix
    burst = 4000
    t0 = time.perf_counter()
    for i in range(burst):
        submit(i)
    srv.drain()
    cap_rate = burst / (time.perf_counter() - t0)
    collect()
    lat.clear()
    lam = rate if rate else util * cap_rate

    # the Poisson stream
    gaps = rng.exponential(1.0 / lam, size=int(lam * seconds * 1.2) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < seconds]
    st0 = srv.stats()
    t0 = time.perf_counter()
    next_collect = t0
    for i, ta in enumerate(arrivals):
        while True:
            now = time.perf_counter()
            if now - t0 >= ta:
                break
            if now >= next_collect:
                collect()
                next_collect = now + 0.002
        submit(i)
    srv.drain()
    t1 = time.perf_counter()
    collect()
    st1 = srv.stats()
    done = len(lat)
    lat_ms = sorted(1e3 * x for x in lat)

    # error of each ring slot's last result against the FP64 path on its input
    errs = {}
    for (key, r), j in last.items():
        B, kind = key
        if len(errs.get(key, [])) >= 4:
            continue
        N = 2 * B
        out = outs[key][r].to(torch.complex128)
        if kind == P.PFS_REQ_INVERSE:
            ref = torch.zeros(N ** 3, dtype=torch.complex128, device="cuda")
            ctx.idsoft_dev(B, inputs[key][j].to(torch.complex128), ref, stream=stream)
        else:
            ref = torch.empty(P.coeff_count(B), dtype=torch.complex128, device="cuda")
            ctx.dsoft_dev(B, inputs[key][j].to(torch.complex128), ref, stream=stream)
        e = (torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref)).item()
        errs.setdefault(key, []).append(e)
    srv.close()
    name = {P.PFS_REQ_FORWARD: "dsoft", P.PFS_REQ_INVERSE: "idsoft"}
    return {
        "bandwidths": SERVE_BANDWIDTHS, "probabilities": SERVE_PROBS, "forward_fraction": 0.5,
        "hbm_budget_bytes": budget_bytes, "slots": slots, "max_batch": max_batch,
        "capacity_per_s": cap_rate, "offered_rate_per_s": lam, "target_utilisation": util if not rate else None,
        "requests": int(len(arrivals)), "completed": done, "seconds": t1 - t0,
        "transforms_per_s": done / (t1 - t0),
        "latency_ms": {"p50": lat_ms[len(lat_ms) // 2], "p99": lat_ms[min(len(lat_ms) - 1, int(0.99 * len(lat_ms)))],
                       "max": lat_ms[-1]},
        "batches": st1["batches"] - st0["batches"],
        "mean_batch": (st1["signals"] - st0["signals"]) / max(1, st1["batches"] - st0["batches"]),
        "max_inflight_batches": st1["max_inflight_batches"], "server_device_bytes": st1["device_bytes"],
        "rel_l2_vs_fp64_max": {"%s_B%d" % (name[k[1]], k[0]): max(v) for k, v in sorted(errs.items())},
    }


def multi_batch_leg(torch, P, ctx, seconds, seed, rate, max_batch=256, burst=2000):
    """The serving mix through the mixed-bandwidth batch entry points
    (pfs_{d,id}soft_batch_multi_f32_dev) on one stream, continuous batching on
    the host: whenever the previous batch is done, every request that has
    arrived (up to max_batch) goes into the next batch, forward and inverse
    requests as one call each.  Reports capacity (a saturating burst), p50/p99
    latency (arrival to batch completion) at the offered rate `rate` -- the
    slot server's offered rate, so the two designs are compared at the same
    load -- the error of every bandwidth and kind against the FP64 path, and
    the fixed-composition batch time with and without graph replay."""
    import numpy as np
    stream = torch.cuda.current_stream()
    rng = np.random.default_rng(seed + 7)
    nin, ring = 4, {16: 512, 32: 512, 64: 160, 128: 64}
    inputs, outs = {}, {}
    for B in SERVE_BANDWIDTHS:
        N, n = 2 * B, P.coeff_count(B)
        c = P.seeded_coeffs_torch(B, [seed + 100 * B + i for i in range(nin)], "cuda", dtype=torch.complex64)
        g = torch.empty((nin, N ** 3), dtype=torch.complex64, device="cuda")
        ctx.idsoft_batch_f32(B, nin, c, g, stream=stream)
        inputs[(B, P.PFS_REQ_INVERSE)], inputs[(B, P.PFS_REQ_FORWARD)] = c, g
        outs[(B, P.PFS_REQ_INVERSE)] = torch.empty((ring[B], N ** 3), dtype=torch.complex64, device="cuda")
        outs[(B, P.PFS_REQ_FORWARD)] = torch.empty((ring[B], n), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    pos = {k: 0 for k in outs}
    last = {}

    def draw():
        return int(rng.choice(SERVE_BANDWIDTHS, p=SERVE_PROBS)), int(rng.integers(0, 2))

    def launch(reqs):
        done_ev = torch.cuda.Event()
        for kind, call in ((P.PFS_REQ_FORWARD, ctx.dsoft_batch_multi), (P.PFS_REQ_INVERSE, ctx.idsoft_batch_multi)):
            sel = [r for r in reqs if r[1] == kind]
            if not sel:
                continue
            bws, ins, os_ = [], [], []
            for (B, k, j) in sel:
                r = pos[(B, k)]
                pos[(B, k)] = (r + 1) % ring[B]
                bws.append(B)
                ins.append(inputs[(B, k)][j])
                os_.append(outs[(B, k)][r])
                last[(B, k)] = (r, j)
            call(bws, ins, os_, stream=stream)
        done_ev.record(stream)
        return done_ev

    ctx.tune("batch_graphs", 0)  # compositions change batch to batch
    # capacity: a saturating burst, batches of max_batch
    reqs = [draw() + (i % nin,) for i in range(burst)]
    t0 = time.perf_counter()
    for b0 in range(0, burst, max_batch):
        launch(reqs[b0:b0 + max_batch])
    torch.cuda.synchronize()
    cap = burst / (time.perf_counter() - t0)
    # the Poisson stream at the offered rate
    lam = rate or 0.7 * cap
    arr = np.cumsum(rng.exponential(1.0 / lam, size=int(lam * seconds * 1.2) + 16))
    arr = arr[arr < seconds]
    kinds = [draw() for _ in range(len(arr))]
    lat, nb, i, inflight = [], 0, 0, None
    t0 = time.perf_counter()
    while i < len(arr) or inflight is not None:
        now = time.perf_counter() - t0
        if inflight is not None and inflight[0].query():
            lat.extend(now - a for a in inflight[1])
            inflight = None
        if inflight is None and i < len(arr) and arr[i] <= now:
            j = i
            while j < len(arr) and arr[j] <= now and j - i < max_batch:
                j += 1
            ev = launch([kinds[k] + (k % nin,) for k in range(i, j)])
            inflight = (ev, list(arr[i:j]))
            nb += 1
            i = j
    t1 = time.perf_counter()
    lat_ms = sorted(1e3 * x for x in lat)
    # error of each bandwidth and kind's last result against the FP64 path
    errs = {}
    name = {P.PFS_REQ_FORWARD: "dsoft", P.PFS_REQ_INVERSE: "idsoft"}
    for (B, kind), (r, j) in sorted(last.items()):
        out = outs[(B, kind)][r].to(torch.complex128)
        if kind == P.PFS_REQ_INVERSE:
            ref = torch.zeros((2 * B) ** 3, dtype=torch.complex128, device="cuda")
            ctx.idsoft_dev(B, inputs[(B, kind)][j].to(torch.complex128), ref, stream=stream)
        else:
            ref = torch.empty(P.coeff_count(B), dtype=torch.complex128, device="cuda")
            ctx.dsoft_dev(B, inputs[(B, kind)][j].to(torch.complex128), ref, stream=stream)
        errs["%s_B%d" % (name[kind], B)] = (torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref)).item()
    # one fixed composition (64 requests of the mix), graph replay against direct launches
    fixed = [draw() + (k % nin,) for k in range(64)]
    fixed_ms = {}
    for graphs in (0, 1):
        ctx.tune("batch_graphs", graphs)
        launch(fixed)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(10):
            launch(fixed)
        torch.cuda.synchronize()
        fixed_ms["graphs" if graphs else "direct"] = (time.perf_counter() - t) / 10 * 1e3
    ctx.tune("batch_graphs", 1)
    return {
        "design": "mixed-bandwidth batch entry points, one stream, host-side continuous batching",
        "max_batch": max_batch, "capacity_per_s": cap, "offered_rate_per_s": lam,
        "requests": int(len(arr)), "completed": len(lat), "seconds": t1 - t0, "batches": nb,
        "transforms_per_s": len(lat) / (t1 - t0),
        "latency_ms": {"p50": lat_ms[len(lat_ms) // 2], "p99": lat_ms[min(len(lat_ms) - 1, int(0.99 * len(lat_ms)))],
                       "max": lat_ms[-1]},
        "rel_l2_vs_fp64": errs,
        "fixed_64_request_batch_ms": fixed_ms,
    }


def dwt_accuracy(ctx, np, max_bw, runs, seed):
    """DWT round-trip accuracy on the device's Wigner-d tables, as the reference's
    benchmark/benchmark_dwt_accuracy.cpp measures it: for BW = 2, 4, ..., max_bw and
    (M, M') = (0, 0), (BW/2, 0), (BW/2, BW/2), random f in [-1, 1] + i[-1, 1], the
    inverse s = d^T f and the forward g = d_w s, and the mean over runs of
    |g - f| / |f| and |g - f|.  Corrected: the reference builds the unweighted
    (0, 0) table into dw at :60, so its dt stays zero there."""
    rng = np.random.default_rng(seed)
    cases = [("M=0,M'=0", lambda bw: (0, 0)), ("M=BW/2,M'=0", lambda bw: (bw // 2, 0)),
             ("M=M'=BW/2", lambda bw: (bw // 2, bw // 2))]
    out = {"cases": [c[0] for c in cases], "runs": runs, "bandwidths": [], "relative": [], "absolute": [],
           "note": "device tables (pfs_wigner_d_matrix); the reference's table mix-up at "
                   "benchmark_dwt_accuracy.cpp:60 corrected"}
    bw = 2
    while bw <= max_bw:
        rel_row, abs_row = [], []
        for _, mm in cases:
            M, Mp = mm(bw)
            dw = ctx.wigner_d_matrix(bw, M, Mp, True)
            dt = ctx.wigner_d_matrix(bw, M, Mp, False)
            r = a = 0.0
            for _ in range(runs):
                f = rng.uniform(-1, 1, dw.shape[0]) + 1j * rng.uniform(-1, 1, dw.shape[0])
                d = float(np.linalg.norm(dw @ (dt.T @ f) - f))
                r += d / float(np.linalg.norm(f))
                a += d
            rel_row.append(r / runs)
            abs_row.append(a / runs)
        out["bandwidths"].append(bw)
        out["relative"].append(rel_row)
        out["absolute"].append(abs_row)
        bw *= 2
    return out


def copy_calibration(torch, P, steps, mib):
    """Copies of known size for scaling PMC byte counters: the library's copy
    kernels with the FFT passes' access widths (16 bytes for complex128, 8 for
    complex64), each with ordinary and with non-temporal loads and stores
    (kernels k_copy16<false> / <true>, k_copy8<false> / <true>)."""
    ctx = P.Context(0)
    n = mib * (1 << 20)
    a = torch.ones(n // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for width in (16, 8):
        for nt in (0, 1):
            for _ in range(steps):
                ctx.copy_dev(b, a, n, nontemporal=nt, width=width)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n, "bytes_written": n, "launches_per_kernel": steps,
                                           "kernels": ["k_copy16<false>", "k_copy16<true>", "k_copy8<false>",
                                                       "k_copy8<true>"]}}))
    ctx.close()


def progress(msg):
    """One line on stderr as each leg of the run ends (stdout keeps only the JSON line)."""
    print("bench.py: " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # steady state: a few round trips leave the GPU clocks still ramping (the
    # same kernels time 7-10 % slower after 3 warm-up steps than after 10)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--bandwidth", type=int, default=128)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--power", type=float, default=0.0,
                    help="coefficient decay exponent s(l) = (l+1)^-power (0 = 1.5 at B >= 256, else 1)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-pairs", type=int, default=3)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    ap.add_argument("--fft-chunk", type=int, default=0, help="beta-slices per FFT chunk (0 = automatic)")
    ap.add_argument("--dwt-variant", type=int, default=0, help="work shape of the FP64 DWT kernels (0 = default)")
    ap.add_argument("--dwt-inv-kernel", type=int, default=-1,
                    help="inverse FP64 DWT kernel: 0 MFMA, 1..3 vector-ALU shapes (-1 = library default)")
    ap.add_argument("--dwt-pair-order", type=int, default=-1,
                    help="pair launch order of the FP64 DWT kernels, bit 0 forward, bit 1 inverse (0 = per-XCD M' "
                         "ranges, 1 = runs of 8 M' per XCD; -1 = library default)")
    ap.add_argument("--batch-signals", type=int, default=64,
                    help="signals per batched mixed-precision round trip at the same bandwidth (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--batch-check", type=int, default=0, help="signals checked against FP64 (0 = all)")
    ap.add_argument("--batch-fft-signals", type=int, default=-1, help="batched FFT chunk: signals (-1 = library default)")
    ap.add_argument("--batch-fft-slices", type=int, default=-1, help="batched FFT chunk: beta-slices (-1 = library default)")
    ap.add_argument("--no-c4", action="store_true", help="skip the 512-signal B=64 batched leg")
    ap.add_argument("--batch-fft-onepass", type=int, default=-1,
                    help="batched FFT at 2B = 64, 128 in one pass (1) or two (0); -1 = library default")
    ap.add_argument("--batch-fft-wide", type=int, default=-1,
                    help="batched FFT passes with 16-byte accesses: 0 none, 1 all, 2 forward passes of 2B <= 128; "
                         "-1 = library default")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched legs (for profiling them); prints no bench line")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="run the Poisson request stream through the request server for this long (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=0.0, help="offered requests/s (0 = 70%% of measured capacity)")
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-max-batch", type=int, default=64)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only (the SO(3) transform of one signal, or a batch, lives "
              "on one device); --gpus %d / WORLD_SIZE=%d is not supported" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import pfsoftlib_amd as P

    if args.copy_calibration:
        copy_calibration(torch, P, max(args.steps, 1), args.copy_calibration)
        return

    B = args.bandwidth
    N = 2 * B
    ctx = P.Context(0)
    ctx.warm(B)
    ctx.tune("fft_chunk_slices", args.fft_chunk)
    ctx.tune("dwt_variant", args.dwt_variant)
    if args.dwt_pair_order >= 0:
        ctx.tune("dwt_pair_order", args.dwt_pair_order)
    if args.dwt_inv_kernel >= 0:
        ctx.tune("dwt_inv_kernel", args.dwt_inv_kernel)
    if args.batch_fft_signals >= 0:
        ctx.tune("batch_fft_signals", args.batch_fft_signals)
    if args.batch_fft_slices >= 0:
        ctx.tune("batch_fft_slices", args.batch_fft_slices)
    if args.batch_fft_onepass >= 0:
        ctx.tune("batch_fft_onepass", args.batch_fft_onepass)
    if args.batch_fft_wide >= 0:
        ctx.tune("batch_fft_wide", args.batch_fft_wide)
    if args.only_batched:
        res = {}
        if args.batch_signals and B <= 128:
            res["batched_mixed"] = batched_leg(torch, P, ctx, B, args.batch_signals, args.batch_steps, 1, args.seed,
                                               args.batch_check)
            torch.cuda.empty_cache()
        if not args.no_c4:
            res["batched_mixed_512xB64"] = batched_leg(torch, P, ctx, 64, 512, args.batch_steps, 1, args.seed,
                                                       args.batch_check)
        print(json.dumps({"only_batched": res}))
        ctx.close()
        return
    power = args.power if args.power > 0 else (1.5 if B >= 256 else 1.0)
    coeffs = torch.from_numpy(P.seeded_coeffs(B, args.seed, power)).cuda()
    grid = torch.zeros(N ** 3, dtype=torch.complex128, device="cuda")
    back = torch.empty_like(coeffs)
    stream = torch.cuda.current_stream()

    def step():
        ctx.idsoft_dev(B, coeffs, grid, stream=stream)
        ctx.dsoft_dev(B, grid, back, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    def kernel_table(timings, steps):
        table = {}
        for name, (cnt, tot_ms) in timings.items():
            per_step_s = tot_ms / steps / 1e3
            ab = algorithmic_bytes_per_step(name, B)
            table[name] = {"launches_per_step": cnt // steps, "us_per_step": round(per_step_s * 1e6, 2),
                           "avg_launch_us": round(tot_ms / cnt * 1e3, 2), "alg_bytes_per_step": ab,
                           "achieved_GBps": round(ab / per_step_s / 1e9, 1),
                           "frac_hbm": round(ab / per_step_s / 1e9 / HBM_PEAK_GBS, 4)}
        return table

    # Per-kernel times come from start/stop events in each timed launch's own
    # dispatch packet.  Each timed launch adds ~3.5 us of device time, so the
    # breakdown of every kernel comes from a pass of its own, outside the timed
    # region, which then times only the kernel with the largest share.
    nbreak = max(10, args.steps // 4)
    ctx.timing(True)
    for _ in range(nbreak):
        step()
    torch.cuda.synchronize()
    ctx.timing(False)
    progress("warm-up and kernel breakdown done (B = %d)" % B)
    kernels = kernel_table(ctx.timing_collect(), nbreak)
    dom = max(kernels, key=lambda k: kernels[k]["us_per_step"])

    # timed region; per-step latency from one event per step boundary (step i
    # runs from marks[i] to marks[i + 1])
    ctx.timing(True, only=[dom])
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    marks[0].record(stream)
    for i in range(args.steps):
        step()
        marks[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.timing(False)
    dk = kernel_table(ctx.timing_collect(), args.steps)[dom]

    lat = sorted(marks[i].elapsed_time(marks[i + 1]) for i in range(args.steps))

    def roofline_line(name, k):
        """The roofline object of the bench line for kernel `name` (timed-region
        figures k): bound, achieved, peak, unit, frac against the binding floor
        (FP64 for the DWT kernels, whose algorithmic flops need longer at
        78.6 TFLOP/s than their bytes at 8 TB/s), the HBM figures beside it,
        counter traffic and the counter-derived issue floor."""
        nl = max(1, k["launches_per_step"])
        ab = k["alg_bytes_per_step"] // nl
        af = algorithmic_flops_per_step(name, B) // nl
        bound, main, hbm = roofline_of(ab, af, k["avg_launch_us"])
        line = {"bound": bound, "kernel": name, "achieved": main["achieved"], "peak": main["peak"],
                "unit": main["unit"], "frac": main["frac"], "traffic": measured_traffic(B, name),
                "alg_bytes": ab, "avg_launch_us": k["avg_launch_us"], "floor_us": main["floor_us"]}
        if af:
            line["alg_flops"] = af
        if hbm:
            line["hbm"] = hbm
        line["issue"] = pipe_report(B, {name: k}).get(name)
        return line

    def kernel_ms(names):
        return round(sum(kernels[k]["us_per_step"] for k in names if k in kernels) / 1e3, 4)

    def fft_stage():
        """Slice-FFT stage per direction: algorithmic bytes (one read and one
        write of the grid, 2 N^3 e) over the time of its two passes, and per
        pass the HBM traffic the committed FETCH_SIZE/WRITE_SIZE counters give
        for that pass over its time (the row/column split moves the grid through
        HBM once per pass)."""
        res = {}
        for d, names in (("forward", ("fft_fwd_rows", "fft_fwd_cols")), ("inverse", ("fft_inv_cols", "fft_inv_rows"))):
            if not all(n in kernels for n in names):
                continue
            us = sum(kernels[n]["us_per_step"] for n in names)
            alg = 2 * N ** 3 * 16
            passes = {}
            for n in names:
                tr = measured_traffic(B, n)
                if tr:
                    gbs = tr / (kernels[n]["avg_launch_us"] * 1e-6) / 1e9
                    passes[n] = {"traffic_bytes": tr, "traffic_GBps": round(gbs, 1),
                                 "traffic_frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
            res[d] = {"us": round(us, 2), "alg_bytes": alg, "alg_GBps": round(alg / (us * 1e-6) / 1e9, 1),
                      "alg_frac_hbm": round(alg / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4), "passes": passes}
        return res

    err = (torch.linalg.vector_norm(back - coeffs) / torch.linalg.vector_norm(coeffs)).item()
    elapsed = t1 - t0
    out = {
        "metric": "B=128 DSOFT+iDSOFT/s: FP64 plain vs batched mixed @≤1e-6 err; p99 latency",
        "value": args.steps / elapsed,
        "unit": "DSOFT+iDSOFT pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: SplitMix64-seeded coefficients, uniform[-1,1]/(l+1)^%g" % power,
        "config": {"workload": "plain FP64 single-signal DSOFT+iDSOFT round trip, B=%d, %d^3 complex128 grid, "
                               "%d coefficients, inputs resident in HBM" % (B, N, P.coeff_count(B)),
                   "bandwidth": B, "grid": [N, N, N], "precision": "fp64", "signals_per_step": 1},
        "roofline": roofline_line(dom, dk),
        "latency_ms": {"p50": lat[len(lat) // 2], "p99": lat[min(len(lat) - 1, int(0.99 * len(lat)))]},
        # kernel time per direction (sum of its kernels' average durations)
        "idsoft_ms": {"kernels_avg": kernel_ms(["dwt_inv", "nyquist_gather", "fft_inv_cols", "fft_inv_rows"])},
        "dsoft_ms": {"kernels_avg": kernel_ms(["fft_fwd_rows", "fft_fwd_cols", "dwt_fwd"])},
        "transforms_per_s": 2 * args.steps / elapsed,
        "roundtrip_rel_l2": err,
        "kernels": kernels,
        "fft_stage": fft_stage(),
        "pipes": pipe_report(B, kernels),
        "kernels_note": "per-kernel breakdown from a separate pass of %d round trips with every kernel timed; "
                        "the timed region times only %s (the roofline kernel)" % (nbreak, dom),
    }
    # PCIe-inclusive rate of the host-buffer entry points (not the headline value)
    c_host = coeffs.cpu().numpy()
    g_host = ctx.idsoft(B, c_host)
    t2 = time.perf_counter()
    for _ in range(3):
        g_host = ctx.idsoft(B, c_host)
        ctx.dsoft(B, g_host)
    out["host_buffer_pairs_per_s"] = 3 / (time.perf_counter() - t2)
    progress("timed region done: %.1f pairs/s; host-buffer leg done" % (args.steps / elapsed))
    del grid, back
    if args.batch_signals and B <= 128:  # the batched path serves B <= 128
        bl = batched_leg(torch, P, ctx, B, args.batch_signals, args.batch_steps, 1, args.seed, args.batch_check)
        out["batched_mixed"] = bl
        out["batched_vs_plain_fp64"] = bl["pairs_per_s"] / out["value"]
        torch.cuda.empty_cache()
        progress("batched leg done")
    if not args.no_c4:
        out["batched_mixed_512xB64"] = batched_leg(torch, P, ctx, 64, 512, args.batch_steps, 1, args.seed, args.batch_check)
        # the split-bf16 contractions on the same workload, with their error against FP64
        for prec in (3, 6):
            r = batched_leg(torch, P, ctx, 64, 512, args.batch_steps, 1, args.seed, args.batch_check, prec)
            out["batched_512xB64_bf16x%d" % prec] = {
                k: r[k] for k in ("precision", "pairs_per_s", "ms_per_batch_roundtrip", "rel_l2_vs_fp64")}
            out["batched_512xB64_bf16x%d" % prec]["kernels"] = {
                k: v for k, v in r["kernels"].items() if k.startswith("batch_dwt_")}
        progress("512 x B = 64 legs done")
    if args.serve_seconds > 0:
        out["serving"] = serving_leg(torch, P, ctx, args.serve_seconds, args.seed, int(args.serve_budget_gb * 2 ** 30),
                                     args.serve_slots, args.serve_max_batch, rate=args.serve_rate or None)
        torch.cuda.empty_cache()
        out["serving_multi_batch"] = multi_batch_leg(torch, P, ctx, args.serve_seconds, args.seed,
                                                     out["serving"]["offered_rate_per_s"])
        progress("serving legs done")
    out["dwt_accuracy"] = dwt_accuracy(ctx, np, 256, 10, args.seed)
    progress("DWT accuracy table done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(B, args.cpu_pairs, args.seed)
        out["cpu_other_configs"] = cpu_other_configs(args.seed)
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
