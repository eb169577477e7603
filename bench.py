"""Benchmark: primes tested/s at p ~ 2e13 (every prime in [2e13, 2e13+1e7)).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--engine batch|session] [--batch-steps S]

A full search of this range applies ~6.24e16 sieving-prime factors (days of
GPU time), so the run is measured in windows and extrapolated in closed form:

* The sieving-prime axis [2, 1e13] has three regions by the types whose
  factorial target is still above q: all three types (q <= T0 = 3.33e12),
  types 1-2 (q <= T1 = 5e12) and type 2 only (q <= T2 = 1e13).
* A step = one segment (SPAN integers) of that axis for all 327,107 test
  primes: sieve + products + reduction. Steps cycle through windows just
  below T0, T1 and T2.
* Each region's rate (factor applications/s, measured over its steps) times
  its factor count (test primes of the active types x sieving primes of the
  region, pi from the logarithmic integral) gives its time; iterate and the
  final check are timed once on top. primes/s = 327,107 / total time.

The same windows are timed on the reference's launch structure (one test
prime per launch pair, the "plain" path) for the batched/plain ratio, and
the 256-range batched workload (starts log-uniform in [1e9, 1e14] from seed
1234, widths uniform in [1e5, 1e7]) is measured the same way at the
quantiles of its factor-application mass.

The roofline line is for the dominant kernel. At 3 limbs that is
k_mfma_runs, the run products as an int8 GEMM on the matrix cores: its
algorithmic int8 multiply-adds per launch (12 byte columns x the bytes of each
128-prime sub-chunk product, per test prime it is applied to) x 2 over its
mean launch time (HIP events on the library's stream around it), against
MI355X's dense int8 peak of 5.0 POP/s. The vector-unit side (k_batch_mul,
which keeps the records across run boundaries) is reported beside it in
30x30-bit partial products against the peak v_mad_u64_u32 issue rate (CUs x
64 lanes x clock).
"""
from __future__ import annotations

import argparse
import bisect
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from clwilson_amd.workloads import batch_ranges, pi_approx  # noqa: E402

METRIC = "primes tested/s at p≈2e13 (batched vs plain), % of int-MAD roofline"
PMIN, PMAX = 20_000_000_000_000, 20_000_010_000_000
SPAN = 4370 * 61440  # integers per step (one default segment)

# HBM traffic of k_mfma_runs on one SPAN segment just below T0 (two launches
# of `bench.py --profile-run` under rocprofv3 --pmc FETCH_SIZE, then --pmc
# WRITE_SIZE; KiB per dispatch). The counters are scaled by what they read
# for a copy of 256 MiB in the same run: FETCH_SIZE reads 131,088 KiB for the
# 262,144 KiB read (x2.0), WRITE_SIZE 262,144 KiB (x1.0). Summary in
# profiles/mfma_runs_cfg3_counters.md. `source_sha256` is the hash of the
# kernel's source file when the counters were taken: if the file has changed
# since, the figure belongs to another kernel and the line reports null.
MFMA_SOURCE = os.path.join(ROOT, "clwilson_amd", "csrc", "clw_mfma_kernels.inc")
MFMA_TRAFFIC = dict(fetch_kib=30703995.2, write_kib=572724.0, fetch_scale=262144 / 131088.0, write_scale=1.0,
                    source_sha256="073f8db48098ebd68dbbf429f92a3c9a513e34461c64ecd62ab398431e616d4f",
                    window="one segment of 268,492,800 sieving integers just below T0 = 3.33e12")
# The reference formulation's declared cost per sieving-prime factor
# application at p ~ 2e13 in 32-bit partial products (mul_r2 22 + square 22.9
# + mul 10.7 + accumulate 26): the same constant for every path of this band.
DECLARED_MAD_PER_FACTOR = 82
SEG_APPS_256 = 10**14  # factor applications per segment of the 256-range batch
INT8_PEAK_OPS = 5.0e15  # MI355X dense int8 matrix-core peak (ops/s; a multiply-add is 2)


def say(*a):
    print("#", *a, file=sys.stderr, flush=True)


def quantile_q(a, b, frac):
    """q in (a, b) with pi(q) - pi(a) = frac (pi(b) - pi(a))."""
    pa, pb = pi_approx(a), pi_approx(b)
    want = pa + frac * (pb - pa)
    lo, hi = float(a), float(b)
    for _ in range(200):
        mid = (lo + hi) / 2
        if pi_approx(mid) < want:
            lo = mid
        else:
            hi = mid
    return int(hi)


# Sieving primes below this are timed in full (one window, no
# extrapolation): there the exponent runs are short, most primes take the
# record path, and the rate is far below the bulk's.
Q_LOW = 10**8


def regions(T, cnt):
    """[(name, qlo, qhi, active types, factor applications)] above Q_LOW."""
    T0, T1, T2 = T
    bounds = [("all types", Q_LOW, T0 + 1, (0, 1, 2)), ("types 1-2", T0 + 1, T1 + 1, (1, 2)),
              ("type 2", T1 + 1, T2 + 1, (2,))]
    out = []
    for name, a, b, types in bounds:
        nprimes = pi_approx(b - 1) - pi_approx(a - 1)
        out.append((name, a, b, types, nprimes * sum(cnt[t] for t in types)))
    return out


class SessionRunner:
    """One range through clw_session_* (lane or wave per test prime)."""

    def __init__(self, cw, dev, mode, per_lane):
        self.cw, self.dev = cw, dev
        self.s = cw.Session(dev, PMIN, PMAX, mode=mode, per_lane=per_lane, segment_span=SPAN)
        self.T, self.cnt, self.limbs = self.s.type_targets, self.s.tpcnt, self.s.limbs
        self.kp = per_lane or int(cw.lib().clw_default_per_lane(self.limbs))
        lane = mode in (0, 1)
        runs = lane and self.limbs <= 4  # lane mode with run products (2-4 limbs)
        self.kernel = (f"k_mul_records_runs<{self.limbs},{self.kp}>" if runs else
                       f"k_mul_records<{self.limbs},{self.kp}>" if lane else f"k_mul_records_lanes<{self.limbs}>")
        self.name = "session, lane per test prime" if lane else "session, wave per test prime"

    def run(self, lo, hi):
        self.s.run(lo, hi)

    def stats(self):
        return self.s.kernel_stats()

    def reset(self):
        self.s.reset_stats()

    def tail(self):
        self.s.iterate()
        self.s.finalize(allow_check_failure=True)  # windows only: rem != 0 expected

    def close(self):
        self.s.close()


class BatchRunner:
    """The range as a one-range batch (clw_batch_*), steps launched eagerly so
    HIP events time each kernel."""

    def __init__(self, cw, dev, per_lane, tune, mfma_slices=0):
        self.cw, self.dev = cw, dev
        # one segment of SPAN integers per window, as the session path runs it
        self.b = cw.Batch(dev, [(PMIN, PMAX)], per_lane=per_lane, segment_span=SPAN, graph=False,
                          segment_apps=10**18, mfma_slices=mfma_slices)
        st = self.b.info[0]
        self.T, self.cnt, self.limbs = list(st.typeTarget), list(st.tpcnt), st.limbs
        # tuned at the top of the all-types region, where the timed windows start
        self.tune = self.b.autotune(max(7, st.typeTarget[0] - 2 * SPAN)) if tune else None
        self.kp = self.tune[0]["per_lane"] if tune else (per_lane or int(cw.lib().clw_batch_default_per_lane(self.limbs)))
        self.vec_kernel = f"k_batch_mul<{self.limbs},{self.kp}>"
        # 3 limbs: the run products go to the matrix cores
        self.kernel = "k_mfma_runs" if self.limbs == 3 else self.vec_kernel
        self.name = "batched engine (one-range batch, persistent work-queue kernel)"
        self._base = None

    def run(self, lo, hi):
        self.b.run(lo, hi)

    def stats(self):
        ks = self.b.kernel_stats()
        if self._base:
            ks = {k: ks[k] - self._base[k] for k in ks}
        return ks

    def reset(self):
        self._base = None
        self._base = self.b.kernel_stats()

    def tail(self):
        self.b.iterate()
        self.b.finalize()

    def close(self):
        self.b.close()


def cpu_baseline(T, cnt, tp, seconds_budget: float) -> dict:
    """The CPU oracle (oracle/wilson_oracle.c, OpenMP) on a bounded sample:
    16 test primes per type over one window per region, scaled by the same
    closed form as the GPU figure."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    cores = min(16, os.cpu_count() or 1)
    types = [0 if int(p) % 3 == 1 else (1 if int(p) % 12 == 5 else 2) for p in tp[:200000:1]]
    sample = {t: [] for t in range(3)}
    for p, t in zip(tp, types):
        if len(sample[t]) < 16:
            sample[t].append(int(p))
    regs = regions(T, cnt)
    per_region = seconds_budget / 3
    rates = []
    for (name, a, b, act, F) in regs:
        ps = [p for t in act for p in sample[t]]
        tg = [T[t] for t in act for p in sample[t]]
        width = 2_000_000
        # grow the window until the sample takes ~per_region seconds
        while True:
            qhi = b - 1
            qlo = qhi - width
            t0 = time.time()
            O.window_product(np.array(ps, dtype=np.uint64), np.array(tg, dtype=np.uint64), qlo, qhi, cores)
            dt = time.time() - t0
            if dt > per_region / 4 or width > 4_000_000_000:
                break
            width *= 4
        apps = len(ps) * pi_approx(qhi) - len(ps) * pi_approx(qlo)
        rates.append((name, apps / dt, dt, qlo, qhi, len(ps)))
    total = sum(F / r[1] for (_, _, _, _, F), r in zip(regs, rates))
    return dict(value=sum(cnt) / total, unit="primes/s", cores=cores, kind="port",
                sample="oracle window products: 16 test primes per active type, one window of "
                       + ", ".join(f"{int(r[4] - r[3]):.3g} integers below {r[0]} boundary ({r[2]:.1f}s)"
                                   for r in rates)
                       + "; factor applications/s scaled by the same closed form as the GPU value",
                factor_apps_per_s={r[0]: r[1] for r in rates})


def measure_plain(dev, cw, regs, wins, steps):
    """The reference's launch structure on the same workload: one test prime
    per launch pair (mullarge + reduce), one window per region (the batched
    leg's first window of the region)."""
    with cw.Session(dev, PMIN, PMAX, mode=3) as s:
        ntp = s.n
        rates = []
        for ri in range(min(steps, 3)):
            lo, hi = wins[ri]
            before = s.kernel_stats()
            t0 = time.perf_counter()
            s.run(lo, hi)
            cw.lib().clw_finish(dev.ctx)
            dt = time.perf_counter() - t0
            apps = s.kernel_stats()["mul_prime_apps"] - before["mul_prime_apps"]
            rates.append(apps / dt)
            say(f"plain region {regs[ri][0]} {dt * 1e3:.1f} ms {apps / dt:.4g} factor apps/s")
        total = sum(F / rates[k] for k, (_, _, _, _, F) in enumerate(regs[:len(rates)]))
        if len(rates) < 3:
            mean = sum(rates) / len(rates)
            total += sum(F / mean for (_, _, _, _, F) in regs[len(rates):])
    return dict(value=ntp / total, region_rates_factor_apps_per_s=rates,
                structure="one test prime per launch pair (k_plain_mul + k_plain_reduce)")


def measure_session(dev, cw, regs, wins, steps):
    """The session path (the reference's kernel structure: one range, one
    launch per type and segment, lane per test prime with run products) on
    one window per region (the batched leg's first), for the
    batched/session ratio."""
    run = SessionRunner(cw, dev, 0, 0)
    rates = []
    try:
        for ri in range(min(steps, 3)):
            lo, hi = wins[ri]
            before = run.stats()
            t0 = time.perf_counter()
            run.run(lo, hi)
            cw.lib().clw_finish(dev.ctx)
            dt = time.perf_counter() - t0
            apps = run.stats()["mul_prime_apps"] - before["mul_prime_apps"]
            rates.append(apps / dt)
            say(f"session region {regs[ri][0]} {dt * 1e3:.1f} ms {apps / dt:.4g} factor apps/s")
        t0 = time.perf_counter()
        run.tail()
        t_tail = time.perf_counter() - t0
    finally:
        run.close()
    total = sum(F / rates[k] for k, (_, _, _, _, F) in enumerate(regs[:len(rates)]))
    if len(rates) < 3:
        mean = sum(rates) / len(rates)
        total += sum(F / mean for (_, _, _, _, F) in regs[len(rates):])
    total += t_tail
    return dict(value=sum(run.cnt) / total, region_rates_factor_apps_per_s=rates, kernel=run.kernel,
                projected_full_search_s=total, structure=run.name)


def mfma_traffic_bytes():
    """Copy-calibrated HBM bytes per k_mfma_runs launch, or None when the
    kernel source differs from the one the counters were taken on."""
    import hashlib
    with open(MFMA_SOURCE, "rb") as f:
        sha = hashlib.sha256(f.read()).hexdigest()
    if not MFMA_TRAFFIC["fetch_kib"] or MFMA_TRAFFIC["source_sha256"] != sha:
        return None
    return (MFMA_TRAFFIC["fetch_kib"] * MFMA_TRAFFIC["fetch_scale"]
            + MFMA_TRAFFIC["write_kib"] * MFMA_TRAFFIC["write_scale"]) * 1024


def measure_batch256(dev, cw, steps: int, tune: bool, target_apps: float = 5e13) -> dict:
    """The 256-range batch: one batch (graph-replayed segment steps), windows
    of the sieving-prime axis at the midpoints of `steps` equal quantiles of
    the factor-application mass F(q) = sum over groups of tp_n x pi(min(q, T));
    each bin's time = its mass / the rate measured in its window."""
    ranges = batch_ranges()
    t0 = time.perf_counter()
    # segments of ~1e14 factor applications (~2 M sieving primes for the 48 M
    # test primes): long enough that the 3-limb band's run products go to the
    # matrix cores (a workgroup's table pays off over ~1,000 sub-chunks)
    b = cw.Batch(dev, ranges, graph=True, segment_apps=SEG_APPS_256)
    setup = time.perf_counter() - t0
    groups = sorted((st.typeTarget[t], st.tpcnt[t]) for st in b.info for t in range(3) if st.tpcnt[t])
    ntp = sum(st.tpcount for st in b.info)
    bands = {}
    for st in b.info:
        bands[st.limbs] = bands.get(st.limbs, 0) + 1
    Ts = [T for T, _ in groups]
    suffix = [0] * (len(groups) + 1)  # test primes of groups k.. (sorted by T)
    for k in range(len(groups) - 1, -1, -1):
        suffix[k] = suffix[k + 1] + groups[k][1]
    below = [0.0] * (len(groups) + 1)
    for k in range(len(groups)):
        below[k + 1] = below[k] + groups[k][1] * pi_approx(Ts[k])

    def F_below(q):
        k = bisect.bisect_left(Ts, q)  # groups with T < q are complete
        return below[k] + suffix[k] * pi_approx(q)

    def active(q):
        return suffix[bisect.bisect_left(Ts, q)]

    F = F_below(Ts[-1])
    say(f"batch256: {len(ranges)} ranges, {ntp} test primes, bands {bands}, F {F:.4g}, setup {setup:.1f}s")

    def quantile(frac):
        lo, hi = 2.0, float(Ts[-1])
        for _ in range(100):
            mid = math.sqrt(lo * hi)
            if F_below(mid) < frac * F:
                lo = mid
            else:
                hi = mid
        return int(hi)

    tune_rep = None
    if tune:
        t0 = time.perf_counter()
        tune_rep = b.autotune(quantile(0.5))
        say(f"batch256 autotune {time.perf_counter() - t0:.1f}s: "
            + "; ".join(f"L={t['limbs']} per_lane {t['per_lane']} items/block {t['items_per_block']}" for t in tune_rep))
    rates, windows = [], []
    for k in range(steps):
        q = quantile((k + 0.5) / steps)
        w = int(target_apps * math.log(q) / max(1, active(q)))
        w = max(61440, min(w, q // 4))
        before = b.kernel_stats()
        t0 = time.perf_counter()
        b.run(q - w, q)
        dt = time.perf_counter() - t0
        ks = b.kernel_stats()
        apps = ks["mul_prime_apps"] - before["mul_prime_apps"]
        segs = ks["segments"] - before["segments"]
        rates.append(apps / dt)
        windows.append(dict(q=q, width=w, active_test_primes=active(q), seconds=dt, factor_apps=apps,
                            segments=segs))
        say(f"batch256 window {k}: [{q - w}, {q}) active {active(q)} {segs} segments {dt * 1e3:.0f} ms "
            f"{apps / dt:.4g} factor apps/s")
    t0 = time.perf_counter()
    b.iterate()
    b.finalize()
    tail = time.perf_counter() - t0
    b.close()
    total = sum(F / steps / r for r in rates) + tail
    return dict(value=ntp / total, unit="primes/s", ranges=len(ranges), test_primes=ntp, limb_bands=bands,
                factor_applications_total=F, projected_full_search_s=total, setup_s=setup, autotune=tune_rep,
                iterate_finalize_s=tail, windows=windows,
                workload="256 disjoint ranges, starts log-uniform in [1e9, 1e14] (seed 1234), widths uniform "
                         "in [1e5, 1e7] rounded to even; one batch, segment steps replayed from a HIP graph")


def measure_serving(seconds: float) -> dict:
    """The request stream (2,000 range workunits, Poisson arrivals of mean
    50/s from seed 99, widths log-uniform in [1e4, 1e7], starts log-uniform
    in [1e9, 1e14]) replayed in real time through the continuous-batching
    request loop under a 64 GiB HBM budget, for a declared wall-clock window.
    Its full work (~7e18 factor applications) takes weeks on one GPU: the
    report counts what completed within the window (latency = completion -
    arrival) and the factor applications/s sustained over it."""
    from clwilson_amd.serve import Server
    from clwilson_amd.workloads import serving_requests
    srv = Server(hbm_budget=64 << 30, streams=3)
    try:
        rep = srv.serve(serving_requests(), duration=seconds, progress=lambda r: say(
            f"serving t={r['wall_s']:.0f}s arrived {r['arrived']} completed {r['completed']} "
            f"in flight {r['in_flight']} queued {r['queued']} {r['factor_apps_per_s']:.3g} factor apps/s"))
    finally:
        srv.close()
    rep["workload"] = ("2,000 requests, Poisson arrivals mean 50/s (seed 99), widths log-uniform [1e4, 1e7], starts "
                       "log-uniform [1e9, 1e14]; 64 GiB HBM budget; window of the stream replayed in real time")
    rep["latency_of"] = "requests completed within the window"
    return rep


def profile_run(cw, args) -> int:
    """A few kernels to read with rocprofv3: the copy of a known size scales
    the FETCH_SIZE / WRITE_SIZE counters to bytes; then k_batch_mul on two
    windows of the production-magnitude range (region boundary T0)."""
    import torch
    n = 256 << 20
    x = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    dev = cw.Device(0)
    run = BatchRunner(cw, dev, args.per_lane, False, args.mfma_slices)
    top = (run.T[0] + 1) // 30 * 30
    for k in range(2):
        run.run(top - (k + 1) * SPAN, top - k * SPAN)
    cw.lib().clw_finish(dev.ctx)
    ks = run.stats()
    run.close()
    dev.close()
    print(json.dumps(dict(copy_bytes=n, kernel=run.kernel, launches=ks["mul_launches"],
                          mul_ms=ks["mul_ms"], mfma_ms=ks["mfma_ms"], record_apps=ks["mul_record_apps"],
                          prime_apps=ks["mul_prime_apps"], partial_products=ks["mul_partial_products"],
                          int8_macs=ks["mfma_int8_macs"], int8_issued=ks["mfma_int8_issued"])))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--engine", choices=["batch", "session"], default="batch",
                    help="batched engine (one-range batch) or the session path")
    ap.add_argument("--mode", type=int, default=0, help="session product kernel (engine session)")
    ap.add_argument("--per-lane", type=int, default=0, help="test primes per lane (0 = library default)")
    ap.add_argument("--mfma-slices", type=int, default=0, help="row slices of k_mfma_runs (0 = library default)")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="windows timed on the one-prime-per-launch path for the batched/plain ratio")
    ap.add_argument("--session-steps", type=int, default=3,
                    help="windows timed on the session path for the batched/session ratio (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=4,
                    help="windows timed on the 256-range batch (0 = skip it)")
    ap.add_argument("--no-tune", action="store_true", help="batches without the autotuner")
    ap.add_argument("--serve-seconds", type=float, default=45.0,
                    help="wall-clock window of the serving replay (0 = skip it)")
    ap.add_argument("--range", type=int, nargs=2, metavar=("PMIN", "PMAX"), default=None,
                    help="measure another range the same way (e.g. 1e11 1e11+1e7 to check the extrapolation "
                         "against a full search); the JSON line then describes that range")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for rocprofv3: a 256 MiB device copy (counter calibration), then two "
                         "one-segment windows of the [2e13, 2e13+1e7) batch (two k_mfma_runs and two k_batch_mul "
                         "launches)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would search its own "
              "ranges, with no collective). Run with --gpus 1.", file=sys.stderr)
        return 2

    import clwilson_amd as cw

    global PMIN, PMAX
    if args.range:
        PMIN, PMAX = args.range
    if args.profile_run:
        return profile_run(cw, args)
    dev = cw.Device(0)
    info = dev.info()
    t_setup = time.time()
    if args.engine == "batch":
        run = BatchRunner(cw, dev, args.per_lane, not args.no_tune, args.mfma_slices)
    else:
        run = SessionRunner(cw, dev, args.mode, args.per_lane)
    T, cnt, L = run.T, run.cnt, run.limbs
    regs = regions(T, cnt)
    # Step windows: region ri (cycling) gets ceil(steps / 3) windows of one
    # segment each, ending at the midpoints of equal quantiles of its factor
    # mass (every active test prime applies every sieving prime of the region,
    # so the mass below q is proportional to pi(q)); the rate varies along a
    # region (exponent runs shorten lower down), so windows stacked at its top
    # alone would overstate it. The plain and session legs (other objects)
    # time the same first window of each region; the warmup uses the top.
    per_region = max(1, -(-args.steps // 3))
    qwin = [[quantile_q(a, b, (j + 0.5) / per_region) // 30 * 30 for j in range(per_region)]
            for (_, a, b, _, _) in regs]
    nxt = [r[2] // 30 * 30 for r in regs]  # warmup windows, from each region's top

    def window(ri):
        hi = nxt[ri]
        lo = hi - SPAN
        nxt[ri] = lo
        return lo, hi

    say(f"setup {time.time() - t_setup:.1f}s: {sum(cnt)} test primes {cnt}, targets {T}, limbs {L}, {run.name}")
    lib = cw.lib()
    for i in range(args.warmup):
        lo, hi = window(i % 3)
        run.run(lo, hi)
    lib.clw_finish(dev.ctx)
    run.reset()
    step_ms = []
    region_time = [0.0, 0.0, 0.0]
    region_apps = [0.0, 0.0, 0.0]
    t_all = time.perf_counter()
    for i in range(args.steps):
        ri = i % 3
        hi = max(qwin[ri][i // 3], regs[ri][1] + SPAN)
        lo = hi - SPAN
        before = run.stats()
        t0 = time.perf_counter()
        run.run(lo, hi)
        lib.clw_finish(dev.ctx)
        dt = time.perf_counter() - t0
        after = run.stats()
        apps = after["mul_prime_apps"] - before["mul_prime_apps"]
        step_ms.append(dt * 1e3)
        region_time[ri] += dt
        region_apps[ri] += apps
        say(f"step {i} region {regs[ri][0]} [{lo}, {hi}) {dt * 1e3:.1f} ms {apps / dt:.4g} factor apps/s")
    wall = time.perf_counter() - t_all
    ks = run.stats()
    # the sieving primes below Q_LOW, in full (a window disjoint from the steps)
    before = run.stats()
    t0 = time.perf_counter()
    run.run(2, Q_LOW)
    lib.clw_finish(dev.ctx)
    t_low = time.perf_counter() - t0
    low_apps = run.stats()["mul_prime_apps"] - before["mul_prime_apps"]
    say(f"sieving primes below {Q_LOW:.0e} in full: {t_low * 1e3:.1f} ms {low_apps / t_low:.4g} factor apps/s")
    # iterate + final check once (their share of a full search)
    t0 = time.perf_counter()
    run.tail()
    t_tail = time.perf_counter() - t0
    run.close()
    first = [(max(qwin[ri][0], regs[ri][1] + SPAN) - SPAN, max(qwin[ri][0], regs[ri][1] + SPAN)) for ri in range(3)]
    plain = measure_plain(dev, cw, regs, first, args.plain_steps) if args.plain_steps > 0 else None
    session = (measure_session(dev, cw, regs, first, args.session_steps)
               if args.session_steps > 0 and args.engine == "batch" else None)

    rates = [region_apps[r] / region_time[r] if region_time[r] > 0 else None for r in range(3)]
    measured = [r for r in rates if r]
    mean_rate = sum(measured) / len(measured)
    total = sum(F / (rates[k] or mean_rate) for k, (_, _, _, _, F) in enumerate(regs)) + t_low + t_tail
    value = sum(cnt) / total

    peak = info["compute_units"] * 64 * info["clock_khz"] * 1e3  # v_mad_u64_u32 / s
    prod_per_rec = 2 * L * L + L
    # algorithmic 30x30-bit partial products counted by the product kernel:
    # 2L^2 + L per full Montgomery product (records, one per run product),
    # L + 1 per 30-bit digit of a run product reduced mod p^2
    vec_ms = ks["mul_ms"] - ks["mfma_ms"]
    achieved = ks["mul_partial_products"] / (vec_ms * 1e-3) if vec_ms > 0 else 0.0
    vec_line = {
        "bound": "valu-int32-mad",
        "kernel": getattr(run, "vec_kernel", run.kernel),
        "achieved": achieved / 1e12,
        "peak": peak / 1e12,
        "unit": "Tmad/s",
        "frac": achieved / peak if peak else None,
        "per_unit": f"30x30-bit partial products per test prime: {L + 1} per 30-bit digit of a run product "
                    f"reduced, {prod_per_rec} per full Montgomery product (records across run boundaries, "
                    "one per run product)",
        "partial_products": ks["mul_partial_products"],
        "launches": ks["mul_launches"],
        "mean_launch_ms": vec_ms / max(1, ks["mul_launches"]),
    }
    if not ks["mul_partial_products"]:
        vec_line["note"] = ("no product on the vector units in these windows: every sub-chunk lies inside one "
                            "exponent run and goes to the matrix cores, so the kernel only walks its record list")
    if ks["mfma_int8_macs"]:
        mfma_launches = max(1, ks["sieve_launches"])
        ops = 2.0 * ks["mfma_int8_macs"] / (ks["mfma_ms"] * 1e-3) if ks["mfma_ms"] else 0.0
        traffic = None
        # the counters were taken with 8 row slices (the untuned default)
        tuned = getattr(run, "tune", None)
        slices = tuned[0].get("mfma_slices", 8) if tuned else (args.mfma_slices or 8)
        if args.range is None and slices == 8 and run.kernel == "k_mfma_runs":
            traffic = mfma_traffic_bytes()
        roofline = {
            "bound": "mfma",
            "kernel": "k_mfma_runs",
            "achieved": ops / 1e12,
            "peak": INT8_PEAK_OPS / 1e12,
            "unit": "TOP/s",
            "frac": ops / INT8_PEAK_OPS,
            "traffic": traffic,
            "traffic_of": MFMA_TRAFFIC["window"] + " (rocprofv3 counters, copy-calibrated)",
            "per_unit": "int8 multiply-adds (2 ops each) per (128-prime sub-chunk inside an exponent run, test "
                        "prime): 12 byte columns x the bytes of the sub-chunk's product (~668 at 2e13: ~8,000)",
            "int8_macs": ks["mfma_int8_macs"],
            "int8_macs_issued": ks["mfma_int8_issued"],
            "launches": mfma_launches,
            "mean_launch_ms": ks["mfma_ms"] / mfma_launches,
            "epilogue_partial_products": ks["mfma_partial_products"],
            "note": "the run products' reduction mod p^2 is linear in the product's bytes, so it runs as an exact "
                    "int8 x int8 -> int32 GEMM on the matrix cores (no floating point, no rounding); the "
                    "vector-unit MAD roofline of the reference formulation is in declared_mad",
            "vector_units": vec_line,
        }
    else:
        roofline = vec_line
    # the reference formulation's MAD count per factor application, the same
    # constant for every path: a rate above the VALU peak measures how much
    # arithmetic the run products and the matrix cores save, not a clock
    F_steps = sum(region_apps)
    t_steps = sum(region_time)
    declared = dict(mad_per_factor_application=DECLARED_MAD_PER_FACTOR,
                    factor_apps_per_s=F_steps / t_steps if t_steps else None,
                    achieved=(F_steps / t_steps * DECLARED_MAD_PER_FACTOR / 1e12) if t_steps else None,
                    peak=peak / 1e12, unit="Tmad/s",
                    frac=(F_steps / t_steps * DECLARED_MAD_PER_FACTOR / peak) if t_steps and peak else None,
                    note="above 1: the reference formulation would need this many 32-bit MADs per second on the "
                         "vector units; the batched engine does ~1/30 of that arithmetic (run products of 128 "
                         "sieving primes reduced once per test prime, the reduction on the matrix cores)")
    batch256 = measure_batch256(dev, cw, args.batch_steps, not args.no_tune) if args.batch_steps > 0 else None
    serving = measure_serving(args.serve_seconds) if args.serve_seconds > 0 else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "primes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sum(step_ms) / len(step_ms),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int (exact: 30-bit limbs in u64 column sums on the vector units, i8 x i8 -> i32 on the "
                 "matrix cores)",
        "data": f"exact: every prime in [{PMIN}, {PMAX}); windows of the sieving-prime axis, extrapolated",
        "config": {
            "workload": ("every prime p in [2e13, 2e13+1e7) (production magnitude, p^2 ~ 2^89)"
                         if (PMIN, PMAX) == (20_000_000_000_000, 20_000_010_000_000)
                         else f"every prime p in [{PMIN}, {PMAX})"),
            "engine": run.name,
            "test_primes": sum(cnt), "types": cnt, "type_targets": T, "limbs": L,
            "step": f"one segment of {SPAN} sieving integers for all test primes",
            "factor_applications_total": sum(r[4] for r in regs) + low_apps,
            "region_rates_factor_apps_per_s": {regs[k][0]: rates[k] for k in range(3)},
            "iterate_finalize_s": t_tail,
            "below_q_low_s": t_low,
            "q_low": Q_LOW,
            "projected_full_search_s": total,
            "per_lane": run.kp,
            "autotune": getattr(run, "tune", None),
        },
        "roofline": roofline,
        "declared_mad": declared,
        "kernel_ms": {k: ks[k] for k in ("sieve_ms", "mfma_ms", "mul_ms", "reduce_ms")},
        "plain": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "session": session,
        "batched_over_session": (value / session["value"]) if session else None,
        "batch256": batch256,
        "serving": serving,
        "wall_s": wall,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(T, cnt, cw.host_primes(PMIN, PMAX), args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
