#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X, through the C ABI.

Main line: stable key-value radix sort (clc_sort) of 2^28 uint32 keys, the
raw outputs of the reference's pcg32 generator seeded srandom(42, 54), with
uint32 values 0..n-1, ascending.

    python bench.py [--steps K] [--warmup W] [--log2n 28] [--key u32|u64]
                    [--legs c1,u64,learned,segmented,batched,primitives | none]
                    [--no-cpu]

A step is one clc_sort call on inputs already resident in HBM. Before each
step the pristine keys/values are restored by a device-to-device copy that
lies outside the step's hipEvent pair, so `value` counts only the sort.
Prints one JSON line: metric, value in GKeys/s, the roofline of the dominant
kernel (onesweep pass), the CPU baseline, and under "legs" the other
workloads of the same path, each timed the same way:

  c1         2^20 pcg32 keys sorted keys-only and an exclusive scan of
             2^20 values pcg32 & 0xFF (Infinity-Cache resident: rates are
             reported, not judged on the HBM roofline);
  u64        2^27 uint64 keys uniform over [0, 2^48) (two pcg32 draws,
             (hi << 32 | lo) masked) + uint32 values;
  learned    clc_learned_sort against clc_sort on 2^28 uint64 keys + uint32
             values from three skewed distributions (Zipf s=1.2 over
             [1, 2^40], round(lognormal(20, 2)), 8 Gaussians with means in
             [0, 2^60) and sd 2^40); outputs compared bit for bit;
  segmented  65,536 segments with lengths round(exp(N(ln 512, 1.5)))
             clamped to [1, 2^20], uint32 keys + segment-local indices:
             one captured clc_graph launch against one clc_sort call per
             segment; outputs compared bit for bit;
  primitives clc_scan (exclusive, "(a + b)" and the run-time compiled
             "b + a + 0"), clc_reduce, clc_copy_if ("(x & 1) == 0"),
             clc_transform ("x * 3u + 1u") and clc_histogram_even (256
             bins) on 2^28 uint32: kernel time (library hipEvent profiler),
             HBM roofline fraction at 8 / 4 / 4 + 4 x kept fraction / 8 / 4
             algorithmic bytes per element, outputs checked by difference /
             int64 sum / torch mask selection / torch arithmetic / bincount;
  batched    clc_learned_segmented_sort against clc_segmented_sort and
             against one clc_sort call per segment on 4096 float32 segments
             with log-uniform sizes 2^10..2^22 (about 2^31 pairs), each from
             uniform[0,1), normal(0,1), exponential(1) or a bimodal normal
             mix with +-0 and duplicates; outputs compared bit for bit.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "opencl-sdk_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
METRIC = "GKeys/s + % HBM roofline, 2^28 u32 pairs; learned-bucket vs plain radix"


def say(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(log2n: int, reps: int) -> dict:
    """The oracle's restatement of the reference sort (4-bit LSD, 8 passes,
    per-thread histograms) on the host cores, on a bounded sample of the
    same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    n = 1 << log2n
    rng = np.random.default_rng(12345)
    keys = rng.integers(0, 1 << 32, size=n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    # The host cores this process may run on, capped by the thread budget
    # the environment grants it (OMP_NUM_THREADS: a one-GPU box's CPU share).
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (
        os.cpu_count() or 1)
    budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = max(1, min(avail, budget) if budget > 0 else avail)
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        oracle.radix_sort(keys, vals, threads=threads)
        times.append(time.perf_counter() - t)
    best = min(times)
    return {"value": round(n / best / 1e9, 4), "unit": "GKeys/s", "cores": threads,
            "kind": "port", "host_cpus": os.cpu_count(), "cpus_available": avail,
            "sample": f"2^{log2n} uniform u32 keys + u32 index values, best of {reps} "
                      f"runs of the oracle's 4-bit LSD restatement "
                      f"({threads} threads of {os.cpu_count()} host CPUs, "
                      f"{best:.2f} s each)"}


PCG_MULT = 6364136223846793005  # pcg32's LCG multiplier (O'Neill, pcg-random.org)


def pcg32_stream(initstate: int, initseq: int, n: int):
    """The first n outputs of pcg32_random_r after pcg32_srandom_r(initstate,
    initseq) (the reference's generator, lib/src/SDK/Random.c:13-50), as a
    numpy uint32 array. The states are produced by doubling: a block of
    states advanced by its own length is one multiply-add with the composed
    LCG step (A, C) -> (A^2, A*C + C); then the output function
    (xorshift-high, random rotation) is applied to every state at once."""
    import numpy as np
    mask = (1 << 64) - 1
    inc = ((initseq << 1) | 1) & mask
    state = (inc + initstate) & mask          # srandom: one step from 0, += initstate
    state = (state * PCG_MULT + inc) & mask   # ... and one more step
    states = np.empty(n, dtype=np.uint64)
    if n == 0:
        return states.astype(np.uint32)
    states[0] = state
    a, c, have = PCG_MULT, inc, 1
    with np.errstate(over="ignore"):
        while have < n:
            take = min(have, n - have)
            np.multiply(states[:take], np.uint64(a), out=states[have:have + take])
            states[have:have + take] += np.uint64(c)
            have += take
            a, c = (a * a) & mask, (c * a + c) & mask
        x = (((states >> np.uint64(18)) ^ states) >> np.uint64(27)).astype(np.uint32)
        rot = (states >> np.uint64(59)).astype(np.uint32)
        del states
        return (x >> rot) | (x << ((np.uint32(32) - rot) & np.uint32(31)))


def event_pairs(torch, k):
    return ([torch.cuda.Event(enable_timing=True) for _ in range(k)],
            [torch.cuda.Event(enable_timing=True) for _ in range(k)])


def check_sorted_pairs(torch, keys0, keys, vals, wide):
    """Ascending keys, values a permutation consistent with the keys, ties
    in index order (stable)."""
    k = keys if wide else keys.to(torch.int64) & 0xFFFFFFFF
    ok = bool((k[1:] >= k[:-1]).all())
    ok = ok and bool((keys0[vals.long()] == keys).all())
    ties = k[1:] == k[:-1]
    return ok and bool((vals[1:][ties] > vals[:-1][ties]).all())


def plain_sort(torch, clc_hip, ctx, wide, log2n, steps, warmup, traffic_dir=None,
               values=True):
    """Times clc_sort on uniform keys (with index values unless values is
    False); returns the measurements."""
    n = 1 << log2n
    kbytes = 8 if wide else 4
    vbytes = 4 if values else 0
    passes = 8 if wide else 4
    executed = 6 if wide else 4  # u64 keys below 2^48: the top two bytes are constant
    import numpy as np
    # Seeded pcg32 streams: u32 keys are raw outputs of srandom(42, 54); u64
    # keys are two draws of srandom(42, 56), (hi << 32 | lo) & (2^48 - 1).
    if wide:
        raw = pcg32_stream(42, 56, 2 * n).astype(np.uint64)
        host_keys = ((raw[0::2] << np.uint64(32)) | raw[1::2]) & np.uint64((1 << 48) - 1)
        del raw
        keys0 = torch.from_numpy(host_keys.view(np.int64)).cuda()
    else:
        host_keys = pcg32_stream(42, 54, n)
        keys0 = torch.from_numpy(host_keys.view(np.int32)).cuda()
    del host_keys
    vals0 = torch.arange(n, dtype=torch.int32, device="cuda") if values else None
    keys = torch.empty_like(keys0)
    vals = torch.empty_like(vals0) if values else None
    kt = clc_hip.U64 if wide else clc_hip.U32

    def restore():
        keys.copy_(keys0)
        if values:
            vals.copy_(vals0)

    for _ in range(warmup):
        restore()
        ctx.sort(keys, vals, key_type=kt)
    torch.cuda.synchronize()

    ctx.reset_timings()
    ctx.set_profiling(True)
    starts, ends = event_pairs(torch, steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        restore()
        starts[i].record()
        ctx.sort(keys, vals, key_type=kt)
        ends[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    step_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms = statistics.fmean(step_ms)
    timings = ctx.kernel_timings()
    if values:
        ok = check_sorted_pairs(torch, keys0, keys, vals, wide)
    else:
        k = keys if wide else keys.to(torch.int64) & 0xFFFFFFFF
        k0 = keys0 if wide else keys0.to(torch.int64) & 0xFFFFFFFF
        ok = bool((k[1:] >= k[:-1]).all()) and bool(torch.equal(torch.sort(k0).values, k))

    kernel = "radix_onesweep_pairs" if values else "radix_onesweep_keys"
    one = timings.get(kernel)
    roofline = None
    if one and one["launches"]:
        # Skipped (constant-digit) passes launch but return at once: average
        # over the executed launches only.
        ran = one["launches"] * executed // passes
        avg_ms = one["total_ms"] / ran
        bytes_per_launch = n * 2 * (kbytes + vbytes)  # read + write of key and value
        achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
        traffic = None
        if traffic_dir is not None and values:
            tpath = os.path.join(traffic_dir, f"onesweep_pairs_{'u64' if wide else 'u32'}"
                                              f"_2p{log2n}_traffic.json")
            if os.path.exists(tpath):
                with open(tpath) as f:
                    traffic = json.load(f).get("hbm_bytes_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "kernel": kernel,
                    "avg_launch_ms": round(avg_ms, 4),
                    "algorithmic_bytes_per_launch": bytes_per_launch}
    sort_bytes = n * (kbytes + passes * 2 * (kbytes + vbytes))  # k + P*2(k+v)
    executed_bytes = n * (kbytes + executed * 2 * (kbytes + vbytes))
    del keys0, vals0, keys, vals
    return {
        "n": n, "kbytes": kbytes, "passes": passes, "executed": executed,
        "ms": ms, "gkeys": n / (ms / 1e3) / 1e9, "ok": ok, "roofline": roofline,
        "sort_roofline_frac": sort_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "executed_pass_roofline_frac": executed_bytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
        "step_ms": step_ms, "wall": wall,
        "kernels": {k2: {"launches": v2["launches"],
                         "avg_ms": round(v2["total_ms"] / max(v2["launches"], 1), 4)}
                    for k2, v2 in timings.items()},
    }


# ---- skewed u64 keys, generated on the GPU ----------------------------------

def zipf_keys(torch, n, s, hi, gen):
    """Zipf(s) ranks over [1, hi] by rejection-inversion (Hoermann and
    Derflinger, 1996): invert H(x) = (x^(1-s) - 1) / (1-s), the integral of
    x^-s, and accept k = round(x) if k - x <= c or u >= H(k + 1/2) - k^-s."""
    a = 1.0 - s

    def H(x):
        return (x ** a - 1.0) / a

    def Hinv(y):
        return (1.0 + y * a) ** (1.0 / a)

    h_lo = H(1.5) - 1.0
    h_hi = H(hi + 0.5)
    c = 2.0 - Hinv(H(2.5) - 2.0 ** -s)
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    todo = torch.arange(n, device="cuda")
    while todo.numel():
        u = h_hi + torch.rand(todo.numel(), dtype=torch.float64, device="cuda",
                              generator=gen) * (h_lo - h_hi)
        x = torch.pow(1.0 + u * a, 1.0 / a)
        k = torch.floor(x + 0.5).clamp_(1.0, float(hi))
        ok = (k - x <= c) | (u >= (torch.pow(k + 0.5, a) - 1.0) / a - torch.pow(k, -s))
        out[todo[ok]] = k[ok].to(torch.int64)
        todo = todo[~ok]
        del u, x, k, ok
    return out


def lognormal_keys(torch, n, mu, sigma, gen):
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
    return torch.exp(x.mul_(sigma).add_(mu)).round_().to(torch.int64)


def gaussian_mixture_keys(torch, n, k, span, sd, gen):
    means = torch.rand(k, dtype=torch.float64, device="cuda", generator=gen) * span
    comp = torch.randint(0, k, (n,), device="cuda", generator=gen)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=gen).mul_(sd)
    x += means[comp]
    del comp
    # clamp to [0, 2^64 - 1]; every mean + 6 sd stays far below 2^63
    return x.clamp_(min=0.0).to(torch.int64)


LEARNED_DISTS = {
    "zipf_s1.2_2^40": lambda torch, n, g: zipf_keys(torch, n, 1.2, 1 << 40, g),
    "lognormal_20_2": lambda torch, n, g: lognormal_keys(torch, n, 20.0, 2.0, g),
    "gauss_mix_8": lambda torch, n, g: gaussian_mixture_keys(torch, n, 8, 2.0 ** 60,
                                                             2.0 ** 40, g),
}


def learned_leg(torch, clc_hip, ctx, log2n, steps, warmup, small_log2n=24):
    """The learned-bucket sort of uint64 keys + uint32 index values against
    the plain radix sort, per skewed distribution, at 2^log2n and (the
    grouping forms compared where 4096 buckets fit LDS) at 2^small_log2n:
    the two-pass form (clc_learned_sort: two 8-bit id passes into up to
    65536 buckets) and the single fused scatter into at most 4096 buckets
    (clc_learned_sort_single_scatter)."""
    kb, vb = 8, 4
    # Algorithmic bytes of the learned form: inference + bucket histogram
    # read (k), one scatter into the buckets (2(k+v)), the LDS finish
    # (2(k+v)): k + 4(k+v). Executed by the two-pass form: the id histogram
    # (k), two id passes and the LDS finish (2(k+v) each): k + 6(k+v). The
    # single scatter executes the count (k), a copy of the input (2(k+v)),
    # the scatter and the finish: also k + 6(k+v).
    lbytes = kb + 4 * (kb + vb)
    xbytes = kb + 6 * (kb + vb)
    out = {"n": 1 << log2n, "key": "u64", "value": "u32 index",
           "algorithmic_bytes_per_pair": {"learned": lbytes, "learned_executed": xbytes,
                                          "plain": kb + 8 * 2 * (kb + vb)}}
    grouping = ("learned_id_histogram", "learned_id_onesweep",
                "learned_single_scatter_count", "learned_single_scatter")

    def case(name, gen, n, seed, full):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        keys0 = gen(torch, n, g)
        vals0 = torch.arange(n, dtype=torch.int32, device="cuda")
        bufs = {lab: (torch.empty_like(keys0), torch.empty_like(vals0))
                for lab in ("learned", "single", "plain")}

        def call(label, stats=False):
            k, v = bufs[label]
            if label == "plain":
                ctx.sort(k, v, key_type=clc_hip.U64)
                return None
            return ctx.learned_sort(k, v, key_type=clc_hip.U64, stats=stats,
                                    single_scatter=label == "single")

        res = {}
        for label in ("learned", "single", "plain"):
            k, v = bufs[label]
            times = []
            for it in range(warmup + steps):
                k.copy_(keys0)
                v.copy_(vals0)
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                call(label)
                e1.record()
                torch.cuda.synchronize()
                if it >= warmup:
                    times.append(e0.elapsed_time(e1))
            res[label] = statistics.fmean(times)
        kp, vp = bufs["plain"]
        same = {lab: bool(torch.equal(bufs[lab][0], kp)) and bool(torch.equal(bufs[lab][1], vp))
                for lab in ("learned", "single")}
        # Kernel times of the grouping stage (library hipEvent profiler), and
        # the model statistics, from one more call of each form.
        ktimes, st = {}, {}
        for label in ("learned", "single"):
            k, v = bufs[label]
            k.copy_(keys0)
            v.copy_(vals0)
            torch.cuda.synchronize()
            ctx.reset_timings()
            ctx.set_profiling(True)
            st[label] = call(label, stats=True)
            torch.cuda.synchronize()
            t = ctx.kernel_timings()
            ctx.set_profiling(False)
            ktimes[label] = {kname: round(t[kname]["total_ms"], 4) for kname in grouping
                             if kname in t}
            same[label] = same[label] and bool(torch.equal(k, kp)) and bool(torch.equal(v, vp))
        r = {
            "learned_ms": round(res["learned"], 4), "plain_ms": round(res["plain"], 4),
            "learned_gkeys": round(n / res["learned"] / 1e6, 3),
            "plain_gkeys": round(n / res["plain"] / 1e6, 3),
            "speedup": round(res["plain"] / res["learned"], 3),
            "bit_identical_to_plain": same["learned"],
            "single_scatter_ms": round(res["single"], 4),
            "single_scatter_speedup": round(res["plain"] / res["single"], 3),
            "single_scatter_bit_identical_to_plain": same["single"],
            "grouping_kernels_ms": ktimes,
            "used_model": st["learned"]["used_model"],
            "num_buckets": st["learned"]["num_buckets"],
            "max_bucket_overfill": round(st["learned"]["max_bucket_overfill"], 3),
            "fallback_fraction": round(st["learned"]["fallback_fraction"], 5),
            "single_scatter_num_buckets": st["single"]["num_buckets"],
            "single_scatter_max_bucket_overfill": round(st["single"]["max_bucket_overfill"], 3),
            "single_scatter_fallback_fraction": round(st["single"]["fallback_fraction"], 5),
        }
        if full:
            r.update({
                "learned_roofline_frac": round(lbytes * n / (res["learned"] / 1e3)
                                               / 1e9 / HBM_PEAK_GBS, 4),
                "learned_executed_bytes_frac": round(xbytes * n / (res["learned"] / 1e3)
                                                     / 1e9 / HBM_PEAK_GBS, 4),
                "sample_size": st["learned"]["sample_size"],
                "sample_fit_error": round(st["learned"]["sample_fit_error"], 5),
                "root_log": st["learned"]["root_log"],
                "distinct_keys_in_first_2^20": int(torch.unique(keys0[: 1 << 20]).numel()),
            })
        del keys0, vals0, bufs
        torch.cuda.empty_cache()
        return r

    for i, (name, gen) in enumerate(LEARNED_DISTS.items()):
        say("learned leg", name, f"2^{log2n}")
        out[name] = case(name, gen, 1 << log2n, 1000 + i, True)
        say(name, out[name])
    if small_log2n:
        small = {"n": 1 << small_log2n}
        for i, (name, gen) in enumerate(LEARNED_DISTS.items()):
            say("learned leg", name, f"2^{small_log2n}")
            small[name] = case(name, gen, 1 << small_log2n, 2000 + i, False)
            say(name, small[name])
        out[f"at_2^{small_log2n}"] = small
    return out


def c5_keys_gpu(torch, off, g):
    """float32 keys of the batched learned workload: segment i drawn from
    uniform[0,1), normal(0,1), exponential(1) or a bimodal normal mix
    (i mod 4), with +0, -0 and repeated values injected; no NaN."""
    total = int(off[-1])
    keys = torch.empty(total, dtype=torch.float32, device="cuda")
    for i in range(len(off) - 1):
        b, e = int(off[i]), int(off[i + 1])
        m = e - b
        kind = i % 4
        if kind == 0:
            x = torch.rand(m, generator=g, device="cuda")
        elif kind == 1:
            x = torch.randn(m, generator=g, device="cuda")
        elif kind == 2:
            x = torch.empty(m, device="cuda").exponential_(1.0, generator=g)
        else:
            pick = torch.rand(m, generator=g, device="cuda") < 0.5
            lo = torch.randn(m, generator=g, device="cuda") * 0.5 - 3.0
            hi = torch.randn(m, generator=g, device="cuda") + 4.0
            x = torch.where(pick, lo, hi)
        x[::37] = 0.0
        x[5::41] = -0.0
        x[7::29] = x[3]
        keys[b:e] = x
    return keys


def learned_batched_leg(torch, clc_hip, ctx, steps, warmup, baselines=True):
    """4096 float32 segments, log-uniform sizes 2^10..2^22 (~2^31 pairs in
    total), u32 values: the batched learned sort against the plain
    segmented sort on the same GPU, bit-identical outputs."""
    import numpy as np
    nseg = 4096
    rng = np.random.default_rng(4096)
    sizes = np.rint(2.0 ** rng.uniform(10.0, 22.0, nseg)).astype(np.int64)
    off = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    total = int(off[-1])
    say("learned batched leg", nseg, "segments,", total, "pairs")
    g = torch.Generator(device="cuda")
    g.manual_seed(4096)
    keys0 = c5_keys_gpu(torch, off, g)
    vals0 = torch.arange(total, dtype=torch.int32, device="cuda")
    off_d = torch.from_numpy(off).cuda()
    kl, vl = torch.empty_like(keys0), torch.empty_like(vals0)
    kp, vp = torch.empty_like(keys0), torch.empty_like(vals0)
    res = {}
    for label, k, v, call in (
            ("learned", kl, vl, lambda: ctx.learned_segmented_sort(
                kl, vl, off_d, key_type=clc_hip.F32, stats=False)),
            ("plain", kp, vp, lambda: ctx.segmented_sort(kp, vp, off_d, key_type=clc_hip.F32))):
        times = []
        for it in range(warmup + steps):
            k.copy_(keys0)
            v.copy_(vals0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            call()
            e1.record()
            torch.cuda.synchronize()
            if it >= warmup:
                times.append(e0.elapsed_time(e1))
        res[label] = statistics.fmean(times)
        say("learned batched", label, round(res[label], 3), "ms")
    same = bool(torch.equal(kl, kp)) and bool(torch.equal(vl, vp))
    # The reference's way: one plain sort call per segment (timed once).
    res["per_segment"] = float("nan")
    if baselines:
        kp.copy_(keys0)
        vp.copy_(vals0)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(nseg):
            b, e = int(off[i]), int(off[i + 1])
            ctx.sort(kp[b:e], vp[b:e], key_type=clc_hip.F32)
        e1.record()
        torch.cuda.synchronize()
        res["per_segment"] = e0.elapsed_time(e1)
        same = same and bool(torch.equal(kl, kp)) and bool(torch.equal(vl, vp))
    kl.copy_(keys0)
    vl.copy_(vals0)
    st = ctx.learned_segmented_sort(kl, vl, off_d, key_type=clc_hip.F32, stats=True)
    torch.cuda.synchronize()
    same = same and bool(torch.equal(kl, kp)) and bool(torch.equal(vl, vp))
    out = {
        "segments": nseg, "pairs": total, "key": "f32", "value": "u32 index",
        "min_len": int(sizes.min()), "max_len": int(sizes.max()),
        "learned_ms": round(res["learned"], 3), "plain_segmented_ms": round(res["plain"], 3),
        "learned_gkeys": round(total / res["learned"] / 1e6, 3),
        "plain_gkeys": round(total / res["plain"] / 1e6, 3),
        "speedup": round(res["plain"] / res["learned"], 3),
        "per_segment_plain_ms": round(res["per_segment"], 3),
        "speedup_vs_per_segment": round(res["per_segment"] / res["learned"], 3),
        "bit_identical_to_plain": same,
        "buckets": st["buckets"], "sample_size": st["sample_size"],
        "fit_rejected": st["fit_rejected"], "max_bucket": st["max_bucket"],
        "fallback_fraction": round(st["fallback_fraction"], 5),
    }
    del keys0, vals0, kl, vl, kp, vp, off_d
    torch.cuda.empty_cache()
    return out


def segmented_leg(torch, clc_hip, ctx, steps, warmup, baselines=True):
    import numpy as np
    nseg = 65536
    rng = np.random.default_rng(2024)
    sizes = np.clip(np.rint(np.exp(rng.normal(math.log(512.0), 1.5, nseg))), 1,
                    1 << 20).astype(np.int64)
    off = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum(sizes, out=off[1:])
    total = int(off[-1])
    say("segmented leg", nseg, "segments,", total, "pairs")
    keys0 = torch.from_numpy(pcg32_stream(42, 57, total).view(np.int32)).cuda()
    off_d = torch.from_numpy(off).cuda()
    seg_of = torch.repeat_interleave(torch.arange(nseg, device="cuda"),
                                     torch.from_numpy(sizes).cuda())
    vals0 = (torch.arange(total, device="cuda") - off_d[seg_of]).to(torch.int32)
    del seg_of
    keys = torch.empty_like(keys0)
    vals = torch.empty_like(vals0)

    keys.copy_(keys0)
    vals.copy_(vals0)
    graph = ctx.segmented_sort_graph(keys, vals, off_d, key_type=clc_hip.U32)

    def timed(fn, reps):
        t = []
        for _ in range(reps):
            keys.copy_(keys0)
            vals.copy_(vals0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            t.append(e0.elapsed_time(e1))
        return t

    timed(graph.launch, warmup)
    graph_ms = statistics.fmean(timed(graph.launch, steps))
    kg, vg = keys.clone(), vals.clone()

    def direct():
        ctx.segmented_sort(keys, vals, off_d, key_type=clc_hip.U32)

    timed(direct, warmup)
    direct_ms = statistics.fmean(timed(direct, steps))
    same_direct = bool(torch.equal(keys, kg)) and bool(torch.equal(vals, vg))

    # One clc_sort call per segment, issued from a host loop over raw pointers.
    L = clc_hip.lib()
    kbase, vbase = keys.data_ptr(), vals.data_ptr()
    handle = ctx.handle
    segs = [(int(o) * 4, int(s)) for o, s in zip(off[:-1], sizes)]

    def per_segment(limit=None):
        for o, s in segs[:limit]:
            st = L.clc_sort(handle, ctypes.c_void_p(kbase + o), clc_hip.U32,
                            ctypes.c_void_p(vbase + o), s, 0)
            if st != 0:
                clc_hip.check(st)

    per_seg_ms = per_seg_wall = float("nan")
    same_per_seg = True
    if baselines:
        timed(lambda: per_segment(4096), 1)
        t0 = time.perf_counter()
        per_seg_ms = timed(per_segment, 1)[0]
        per_seg_wall = time.perf_counter() - t0
        same_per_seg = bool(torch.equal(keys, kg)) and bool(torch.equal(vals, vg))
    graph.close()
    # Algorithmic bytes by segment (8-byte pairs): an LDS-resident segment
    # reads and writes its pairs once, 2(k+v); a segment the size-class
    # table sends to the onesweep path costs the plain sort's k + P*2(k+v)
    # with P = 4 digit places; one-key segments move nothing.
    table = ctx.segmented_sort_table(clc_hip.U32, True)
    large_class = 7
    buckets = np.minimum(np.ceil(np.log2(np.maximum(sizes, 2))).astype(np.int64), 31)
    on_large = np.array([table[b] == large_class for b in buckets]) & (sizes >= 2)
    lds_pairs = int(sizes[(sizes >= 2) & ~on_large].sum())
    large_pairs = int(sizes[on_large].sum())
    mixed = lds_pairs * 2 * (4 + 4) + large_pairs * (4 + 4 * 2 * (4 + 4))
    algo = total * 2 * (4 + 4)
    out = {
        "segments": nseg, "pairs": total, "min_len": int(sizes.min()),
        "median_len": float(np.median(sizes)), "max_len": int(sizes.max()),
        "graph_ms": round(graph_ms, 4), "direct_call_ms": round(direct_ms, 4),
        "per_segment_launches_ms": round(per_seg_ms, 2),
        "per_segment_wall_s": round(per_seg_wall, 3),
        "speedup_vs_per_segment": round(per_seg_ms / graph_ms, 1),
        "graph_gkeys": round(total / graph_ms / 1e6, 3),
        "graph_roofline_frac": round(mixed / (graph_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "algorithmic_bytes": mixed,
        "pairs_in_lds_segments": lds_pairs, "pairs_in_onesweep_segments": large_pairs,
        "segments_on_onesweep": int(on_large.sum()),
        "graph_single_pass_frac": round(algo / (graph_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "bit_identical": same_direct and same_per_seg,
    }
    say("segmented", out)
    del keys0, vals0, keys, vals, kg, vg, off_d
    torch.cuda.empty_cache()
    return out


def c1_leg(torch, clc_hip, ctx, steps, warmup):
    """The CPU-runnable configuration: 2^20 raw pcg32 outputs (srandom(42,
    54)) sorted keys-only ascending, and an exclusive plus-scan of 2^20
    values pcg32 & 0xFF (srandom(42, 55)). Each call timed by hipEvents on
    resident inputs (restore copies outside the events). 4 MiB in and out
    stay in the 256 MB Infinity Cache, so the rates are reported but not
    judged against the HBM roofline."""
    import numpy as np
    n = 1 << 20
    keys0 = torch.from_numpy(pcg32_stream(42, 54, n).view(np.int32)).cuda()
    x = torch.from_numpy((pcg32_stream(42, 55, n) & np.uint32(0xFF)).view(np.int32)).cuda()
    keys = torch.empty_like(keys0)
    out = torch.empty_like(x)

    def timed(before, call):
        t = []
        for it in range(warmup + steps):
            before()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            call()
            e1.record()
            torch.cuda.synchronize()
            if it >= warmup:
                t.append(e0.elapsed_time(e1))
        return statistics.fmean(t)

    sort_ms = timed(lambda: keys.copy_(keys0),
                    lambda: ctx.sort(keys, None, key_type=clc_hip.U32))
    k = keys.to(torch.int64) & 0xFFFFFFFF
    sort_ok = bool(torch.equal(torch.sort(keys0.to(torch.int64) & 0xFFFFFFFF).values, k))
    scan_ms = timed(lambda: None,
                    lambda: ctx.scan(x, out, "(a + b)", "0", True, type_code=clc_hip.U32))
    d = out[1:].long() - out[:-1].long()
    scan_ok = int(out[0].item()) == 0 and bool(torch.equal(d, x[:-1].long()))
    res = {
        "n": n, "note": "inputs and outputs (4 MiB each) fit the Infinity Cache: "
                        "reported, not judged on the HBM roofline",
        "sort_keys_only": {"ms": round(sort_ms, 4),
                           "gkeys": round(n / sort_ms / 1e6, 3),
                           "gbs_at_36B_per_key": round(36 * n / sort_ms / 1e6, 1),
                           "output_check": sort_ok},
        "exclusive_scan": {"ms": round(scan_ms, 4),
                           "gelem_s": round(n / scan_ms / 1e6, 3),
                           "gbs_at_8B_per_elem": round(8 * n / scan_ms / 1e6, 1),
                           "output_check": scan_ok},
    }
    say("c1", res)
    return res


def primitives_leg(torch, clc_hip, ctx, steps, warmup):
    """clc_scan / clc_reduce / clc_histogram_even on 2^28 uint32 keys uniform
    over 32 bits. Kernel times come from the library's per-launch hipEvent
    profiler (average over `steps` calls after `warmup`)."""
    n = 1 << 28
    g = torch.Generator(device="cuda")
    g.manual_seed(2024)
    x = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device="cuda",
                      generator=g)
    out = torch.empty_like(x)
    hist = torch.zeros(256, dtype=torch.int32, device="cuda")
    mask = 0xFFFFFFFF

    def timed(fn, names):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ctx.reset_timings()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        t = ctx.kernel_timings()
        ctx.set_profiling(False)
        ms = {k: t[k]["total_ms"] / t[k]["launches"] for k in names if k in t}
        return ms, wall

    def rate(ms, bytes_per):
        gbs = bytes_per * n / (ms * 1e6)
        return {"kernel_ms": round(ms, 4), "gelem_s": round(n / (ms * 1e6), 2),
                "achieved_gbs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}

    res = {"n": n, "dtype": "u32"}
    for label, op, kname in (("scan_exclusive_plus", "(a + b)", "scan"),
                             ("scan_exclusive_rtc_expr", "b + a + 0", "rtc_scan")):
        say("leg primitives", label)
        ms, wall = timed(lambda: ctx.scan(x, out, op, "0", True, type_code=clc_hip.U32),
                         [kname])
        d = (out[1:].long() - out[:-1].long()) & mask
        ok = bool(out[0].item() == 0) and bool(torch.equal(d, x[:-1].long() & mask))
        res[label] = dict(rate(ms[kname], 8), call_ms=round(wall, 4), output_check=ok)
    say("leg primitives reduce")
    want = int(x.long().sum().item()) & mask
    got = []
    ms, wall = timed(lambda: got.append(ctx.reduce(x, "(a + b)", "0", 0,
                                                   type_code=clc_hip.U32)),
                     ["reduce", "reduce_partials"])
    res["reduce_plus"] = dict(rate(ms["reduce"] + ms.get("reduce_partials", 0.0), 4),
                              call_ms_incl_host_sync=round(wall, 4),
                              output_check=all(int(v) == want for v in got))
    say("leg primitives copy_if")
    counts = []
    ms, wall = timed(lambda: counts.append(ctx.copy_if(x, out, "(x & 1) == 0",
                                                       type_code=clc_hip.U32)),
                     ["rtc_copy_if"])
    want = x[(x & 1) == 0]
    kept = want.numel()
    ok = all(c == kept for c in counts) and bool(torch.equal(out[:kept], want))
    del want
    # Algorithmic bytes: read every element, write the kept ones.
    res["copy_if_even"] = dict(rate(ms["rtc_copy_if"], 4 + 4 * kept / n),
                               kept=kept, call_ms_incl_host_sync=round(wall, 4),
                               output_check=ok)
    say("leg primitives transform")
    out2 = torch.empty_like(x)
    ms, wall = timed(lambda: ctx.transform(x, out2, "x * 3u + 1u", in_type=clc_hip.U32,
                                           out_type=clc_hip.U32), ["rtc_transform"])
    ok = bool(torch.equal(out2, x * 3 + 1))  # int32 arithmetic wraps as uint32
    del out2
    res["transform_u32"] = dict(rate(ms["rtc_transform"], 8), call_ms=round(wall, 4),
                                output_check=ok)
    say("leg primitives histogram_even")
    ms, wall = timed(lambda: (hist.zero_(), ctx.histogram_even(
        x, hist, 0.0, 4294967296.0, type_code=clc_hip.U32)), ["histogram_even"])
    # The reference bins in float: x rounds to float first (values within
    # 128 of 2^32 round to 2^32 and fall outside [0, 2^32)); scale 2^-24 is
    # exact.
    xf = (x.long() & mask).to(torch.float32)
    inside = xf < 4294967296.0
    bins = (xf[inside] * (1.0 / 16777216.0)).long()
    ok = bool(torch.equal(hist.long(), torch.bincount(bins, minlength=256)))
    res["histogram_even_256"] = dict(rate(ms["histogram_even"], 4), call_ms=round(wall, 4),
                                     output_check=ok)
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=None,
                    help="log2 of the key count (default 28 for u32, 27 for u64)")
    ap.add_argument("--key", choices=["u32", "u64"], default="u32",
                    help="u32: full 32-bit keys; u64: keys uniform over [0, 2^48)")
    ap.add_argument("--legs", default="c1,u64,learned,segmented,batched,primitives",
                    help="extra workloads to report: comma list of c1, u64, learned, "
                         "segmented, batched, primitives, or 'none'")
    ap.add_argument("--keys-only", action="store_true",
                    help="sort keys without values (main line only)")
    ap.add_argument("--learned-log2n", type=int, default=28)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-baselines", action="store_true",
                    help="skip the one-sort-call-per-segment baselines of the segmented "
                         "and batched legs (for profiler runs, which serialise every "
                         "dispatch)")
    ap.add_argument("--cpu-log2n", type=int, default=28)
    ap.add_argument("--cpu-reps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark runs on exactly one GPU (asked for {args.gpus}); "
              "the sort has no multi-GPU path in this build", file=sys.stderr)
        return 2
    legs = [] if args.legs == "none" else [x for x in args.legs.split(",") if x]
    for x in legs:
        if x not in ("c1", "u64", "learned", "segmented", "batched", "primitives"):
            ap.error(f"unknown leg {x}")

    import torch
    import clc_hip

    wide = args.key == "u64"
    log2n = args.log2n if args.log2n is not None else (27 if wide else 28)
    torch.cuda.set_device(0)
    ctx = clc_hip.Context()

    say("main", args.key, f"2^{log2n}")
    m = plain_sort(torch, clc_hip, ctx, wide, log2n, args.steps, args.warmup,
                   traffic_dir=os.path.join(ROOT, "profiles"), values=not args.keys_only)
    ok = m["ok"]
    step_ms = m["step_ms"]
    line = {
        "metric": METRIC,
        "value": round(m["gkeys"], 3),
        "unit": "GKeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.key,
        "data": "synthetic (seeded pcg32 stream, srandom(42, 54); values 0..n-1)",
        "config": {"workload": (f"stable {'keys-only' if args.keys_only else 'key-value'} "
                                f"radix sort, 2^{log2n} "
                                + ("uint64 keys uniform over [0, 2^48)" if wide else
                                   "uint32 keys uniform over 32 bits")
                                + ("" if args.keys_only else " + uint32 index values")
                                + ", ascending"),
                   "n": m["n"], "key_bytes": m["kbytes"],
                   "value_bytes": 0 if args.keys_only else 4,
                   "passes": m["passes"], "executed_passes": m["executed"],
                   "parallelism": "single GPU"},
        "roofline": m["roofline"],
        "cpu_baseline": None,
        "extra": {"sort_roofline_frac": round(m["sort_roofline_frac"], 4),
                  "executed_pass_roofline_frac": round(m["executed_pass_roofline_frac"], 4),
                  "step_ms_min": round(min(step_ms), 4),
                  "step_ms_median": round(statistics.median(step_ms), 4),
                  "step_ms_max": round(max(step_ms), 4),
                  "wall_s_timed_loop_incl_restore_copies": round(m["wall"], 3),
                  "output_check": ok,
                  "kernels": m["kernels"]},
        "legs": {},
    }
    torch.cuda.empty_cache()
    if "c1" in legs:
        c1 = c1_leg(torch, clc_hip, ctx, max(10, args.steps), 3)
        ok = ok and c1["sort_keys_only"]["output_check"] and c1["exclusive_scan"]["output_check"]
        line["legs"]["c1_2^20"] = c1
    if "u64" in legs and not wide:
        say("leg u64 2^27")
        u = plain_sort(torch, clc_hip, ctx, True, 27, args.steps, args.warmup,
                       traffic_dir=os.path.join(ROOT, "profiles"))
        ok = ok and u["ok"]
        line["legs"]["u64_2^27_uniform48"] = {
            "gkeys": round(u["gkeys"], 3), "ms": round(u["ms"], 4), "output_check": u["ok"],
            "executed_passes": u["executed"], "roofline": u["roofline"],
            "sort_roofline_frac": round(u["sort_roofline_frac"], 4),
            "executed_pass_roofline_frac": round(u["executed_pass_roofline_frac"], 4)}
        torch.cuda.empty_cache()
    if "learned" in legs:
        lr = learned_leg(torch, clc_hip, ctx, args.learned_log2n, max(3, args.steps // 4),
                         1)
        cases = [v for v in lr.values() if isinstance(v, dict) and "speedup" in v]
        cases += [v for sub in lr.values() if isinstance(sub, dict)
                  for v in sub.values() if isinstance(v, dict) and "speedup" in v]
        ok = ok and all(v["bit_identical_to_plain"] and v["single_scatter_bit_identical_to_plain"]
                        for v in cases)
        line["legs"]["learned"] = lr
    if "segmented" in legs:
        sg = segmented_leg(torch, clc_hip, ctx, max(3, args.steps // 2), 2,
                           baselines=not args.no_baselines)
        ok = ok and sg["bit_identical"]
        line["legs"]["segmented"] = sg
    if "batched" in legs:
        lb = learned_batched_leg(torch, clc_hip, ctx, 3, 1, baselines=not args.no_baselines)
        ok = ok and lb["bit_identical_to_plain"]
        line["legs"]["learned_batched"] = lb
    if "primitives" in legs:
        pr = primitives_leg(torch, clc_hip, ctx, max(5, args.steps // 2), 2)
        ok = ok and all(v["output_check"] for v in pr.values() if isinstance(v, dict))
        line["legs"]["primitives"] = pr
        torch.cuda.empty_cache()
    line["extra"]["output_check"] = ok
    if not args.no_cpu:
        say("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(args.cpu_log2n, args.cpu_reps)
    print(json.dumps(line), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
