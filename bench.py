"""Benchmark: primes tested/s at p ~ 2e13 (every prime in [2e13, 2e13+1e7)).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A full search of this range applies ~6.24e16 sieving-prime factors (days of
GPU time), so the run is measured in windows and extrapolated in closed form:

* The sieving-prime axis [2, 1e13] has three regions by the types whose
  factorial target is still above q: all three types (q <= T0 = 3.33e12),
  types 1-2 (q <= T1 = 5e12) and type 2 only (q <= T2 = 1e13).
* A step = one segment (clw_session_run over SPAN integers) of that axis for
  all 327,107 test primes: sieve + products + side powers + reduction.
  Steps cycle through windows just below T0, T1 and T2.
* Each region's rate (factor applications/s, measured over its steps) times
  its factor count (test primes of the active types x sieving primes of the
  region, pi from the logarithmic integral) gives its time; iterate and the
  final check are timed once on top. primes/s = 327,107 / total time.

The roofline line is for the dominant kernel (the record products): the
algorithmic 30x30-bit partial products per launch (21 per record per test
prime with 3 limbs) over the mean launch time from HIP events, against
gfx950's peak v_mad_u64_u32 issue rate (CUs x 64 lanes x clock).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "primes tested/s at p≈2e13 (batched vs plain), % of int-MAD roofline"
PMIN, PMAX = 20_000_000_000_000, 20_000_010_000_000
SPAN = 4370 * 61440  # integers per step (one default segment)


def li(x: float) -> float:
    """Logarithmic integral Li(x) = li(x) - li(2) by Ramanujan's series."""
    g = 0.5772156649015329
    lx = math.log(x)
    s, term, k = 0.0, 1.0, 1
    inner = 0.0
    while True:
        term *= lx / k
        if (k - 1) % 2 == 0:
            inner += 1.0 / (2 * ((k - 1) // 2) + 1)
        add = term * ((-1) ** (k - 1)) / (2 ** (k - 1)) * inner
        s += add
        if abs(add) < 1e-17 * abs(s) and k > 10:
            break
        k += 1
    return g + math.log(lx) + math.sqrt(x) * s - 1.045163780117492784


def pi_approx(x: float) -> float:
    """pi(x) ~ Li(x) - Li(sqrt x)/2 (Riemann's leading terms); relative error
    ~1e-7 at 1e12-1e13, far below the timing noise."""
    return li(x) - 0.5 * li(math.sqrt(x))


def regions(T, cnt):
    """[(name, qlo, qhi, active types, factor applications)]"""
    T0, T1, T2 = T
    bounds = [("all types", 2, T0 + 1, (0, 1, 2)), ("types 1-2", T0 + 1, T1 + 1, (1, 2)),
              ("type 2", T1 + 1, T2 + 1, (2,))]
    out = []
    for name, a, b, types in bounds:
        nprimes = pi_approx(b - 1) - (pi_approx(a - 1) if a > 2 else 0.0)
        out.append((name, a, b, types, nprimes * sum(cnt[t] for t in types)))
    return out


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


def measure_plain(dev, cw, regs, nxt, steps):
    """The reference's launch structure on the same workload: one test prime
    per launch pair (mullarge + reduce), one window per region."""
    out = {}
    with cw.Session(dev, PMIN, PMAX, mode=3) as s:
        rates = []
        for ri in range(min(steps, 3)):
            hi = nxt[ri]
            lo = hi - SPAN
            nxt[ri] = lo
            before = s.kernel_stats()
            t0 = time.perf_counter()
            s.run(lo, hi)
            cw.lib().clw_finish(dev.ctx)
            dt = time.perf_counter() - t0
            apps = s.kernel_stats()["mul_prime_apps"] - before["mul_prime_apps"]
            rates.append(apps / dt)
            print(f"# plain region {regs[ri][0]} {dt * 1e3:.1f} ms {apps / dt:.4g} factor apps/s",
                  file=sys.stderr, flush=True)
        total = sum(F / rates[k] for k, (_, _, _, _, F) in enumerate(regs[:len(rates)]))
        if len(rates) < 3:
            mean = sum(rates) / len(rates)
            total += sum(F / mean for (_, _, _, _, F) in regs[len(rates):])
        out = dict(value=327107 / total, region_rates_factor_apps_per_s=rates,
                   structure="one test prime per launch pair (k_plain_mul + k_plain_reduce)")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--mode", type=int, default=0)
    ap.add_argument("--per-lane", type=int, default=0, help="test primes per lane (0 = library default)")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="windows timed on the one-prime-per-launch path for the batched/plain ratio")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would search its own "
              "ranges; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import clwilson_amd as cw

    dev = cw.Device(0)
    info = dev.info()
    t_setup = time.time()
    sess = cw.Session(dev, PMIN, PMAX, mode=args.mode, per_lane=args.per_lane)
    T = sess.type_targets
    cnt = sess.tpcnt
    regs = regions(T, cnt)
    # windows just below each region's upper end, disjoint across steps
    nxt = [r[2] for r in regs]

    def window(ri):
        hi = nxt[ri]
        lo = hi - SPAN
        nxt[ri] = lo
        return lo, hi

    print(f"# setup {time.time() - t_setup:.1f}s: {sess.n} test primes {cnt}, targets {T}, limbs {sess.limbs}",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        lo, hi = window(i % 3)
        sess.run(lo, hi)
    sess.reset_stats()
    step_ms = []
    region_time = [0.0, 0.0, 0.0]
    region_apps = [0.0, 0.0, 0.0]
    mul_ms = mul_launch = rec_apps = 0.0
    lib = cw.lib()
    lib.clw_finish(dev.ctx)
    t_all = time.perf_counter()
    for i in range(args.steps):
        ri = i % 3
        lo, hi = window(ri)
        before = sess.kernel_stats()
        t0 = time.perf_counter()
        sess.run(lo, hi)
        lib.clw_finish(dev.ctx)
        dt = time.perf_counter() - t0
        after = sess.kernel_stats()
        step_ms.append(dt * 1e3)
        region_time[ri] += dt
        region_apps[ri] += after["mul_prime_apps"] - before["mul_prime_apps"]
        print(f"# step {i} region {regs[ri][0]} [{lo}, {hi}) {dt * 1e3:.1f} ms "
              f"{(after['mul_prime_apps'] - before['mul_prime_apps']) / dt:.4g} factor apps/s",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t_all
    ks = sess.kernel_stats()
    plain = None
    if args.plain_steps > 0:
        plain = measure_plain(dev, cw, regs, nxt, args.plain_steps)
    # iterate + final check once (their share of a full search)
    t0 = time.perf_counter()
    sess.iterate()
    res = sess.finalize(allow_check_failure=True)  # windows only: rem != 0 expected
    t_tail = time.perf_counter() - t0
    sess.close()

    rates = [region_apps[r] / region_time[r] if region_time[r] > 0 else None for r in range(3)]
    measured = [r for r in rates if r]
    mean_rate = sum(measured) / len(measured)
    total = sum(F / (rates[k] or mean_rate) for k, (_, _, _, _, F) in enumerate(regs)) + t_tail
    value = sum(cnt) / total

    peak = info["compute_units"] * 64 * info["clock_khz"] * 1e3  # v_mad_u64_u32 / s
    L = res.limbs
    prod_per_rec = 2 * L * L + L
    achieved = ks["mul_record_apps"] * prod_per_rec / (ks["mul_ms"] * 1e-3) if ks["mul_ms"] else 0.0
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
        "dtype": "u32",
        "data": "exact: every prime in [2e13, 2e13+1e7); windows of the sieving-prime axis, extrapolated",
        "config": {
            "workload": "config 3, plain method: every prime p in [2e13, 2e13+1e7), one range per launch",
            "test_primes": sum(cnt), "types": cnt, "type_targets": T, "limbs": L,
            "step": f"one segment of {SPAN} sieving integers for all test primes",
            "factor_applications_total": sum(r[4] for r in regs),
            "region_rates_factor_apps_per_s": {regs[k][0]: rates[k] for k in range(3)},
            "iterate_finalize_s": t_tail,
            "projected_full_search_s": total,
            "mode": "lane per test prime" if args.mode in (0, 1) else "wave per test prime",
            "per_lane": args.per_lane or int(cw.lib().clw_default_per_lane(L)),
        },
        "roofline": {
            "bound": "valu-int32-mad",
            "kernel": f"k_mul_records<{L}>",
            "achieved": achieved / 1e12,
            "peak": peak / 1e12,
            "unit": "Tmad/s",
            "frac": achieved / peak if peak else None,
            "traffic": None,
            "per_unit": f"{prod_per_rec} 30x30-bit partial products per record per test prime",
            "launches": ks["mul_launches"],
            "mean_launch_ms": ks["mul_ms"] / max(1, ks["mul_launches"]),
        },
        "kernel_ms": {k: ks[k] for k in ("sieve_ms", "mul_ms", "side_ms", "reduce_ms")},
        "plain": plain,
        "batched_over_plain": (value / plain["value"]) if plain else None,
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
