"""Writes the golden fixtures of tests/golden/ with the CPU oracle.

    python tests/golden/make_golden.py [--cfg1] [--direct]
    python tests/golden/make_golden.py --mfma | --cfg2 | --cfg4   (one fixture each)

Fixtures (JSON):
  kat_selftest.json   the reference's seven self-test answers (cl_wilson.cpp:2007-2123)
  cgh_tables.json     Costa-Gerbicz-Harvey Tables 1-2 (w_p for |w_p| small)
  oracle_small.json   per-prime quotients of small ranges computed by the oracle
  windows.json        canonical window products at config-2/3 magnitudes
  cfg1.json           (--cfg1) the whole of [1e6, 2e6): checksum, results, sha256
                      of the Wilson quotient vector (little-endian uint64)
  direct_lt_2e6.json  (--direct) every prime p < 2e6 by the definition,
                      (p-1)! = 2 * 3 * ... * (p-1) mod p^2: per range [5, 1e6)
                      and [1e6, 2e6) the count, checksum, sha256 of the
                      quotient vector, and every 500th (p, quot) as samples
  mfma_windows.json   (--mfma) window products at p ~ 2e13 and at the top of
                      the 3-limb band (2.4876e13), each window ~1.4 M sieving
                      primes (long enough for the matrix-core run products)
                      across exponent-run boundaries T/2, T/3 and targets
  cfg2_oracle.json    (--cfg2) the full plain method for 16 sampled primes of
                      [1e11, 1e11+1e7): product of every sieving prime up to
                      the type target, iterate, final check -> quot
  cfg4_windows.json   (--cfg4) window products of sampled test primes of the
                      256-range batch (workloads.batch_ranges(), seed 1234)
                      over two windows of sieving primes
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

# cl_wilson.cpp:2007-2123 -- (pmin, pmax, results, checksum, pi(maxT), last p, last w)
KATS = [
    (1239053554603, 1239053554604, 1, 0x00000240FAB1A752, 8257082014, 1239053554603, -4),
    (1108967825921, 1108967825922, 1, 0x0000010233A2220D, 10956003002, 1108967825921, 12),
    (5609877309359, 5609877309360, 1, 0x00000A344D7D0F58, 101542897873, 5609877309359, -6),
    (16556218163369, 16556218163370, 1, 0x00000F0ECB80A0AB, 147755473426, 16556218163369, 2),
    (200, 564, 57, 0x00000000000080A3, 30, 563, 0),
    (86000000, 87467200, 1, 0x0000097C61AB0943, 2604536, 87467099, -2),
    (17524177394450, 17524177394618, 1, 0x00005B54B4CBBC47, 304620766446, 17524177394617, 256),
]

# "GERBICZ a search for wilson primes.pdf" p.3073, Tables 1-2 (w_p in (-p/2, p/2)).
CGH_T1 = [(56151923, -1), (4036677373, -5), (11774118061, -1), (5609877309359, -6), (14296621, 2),
          (10746881, -7), (87467099, -2), (11892977, -7), (16556218163369, 2), (39198017, -7),
          (8315831, 3), (1767839071, 8), (93559087, -3), (29085907, 9), (51802061, 4),
          (67133912011, 9), (258818504023, 4), (42647052491, 10), (1239053554603, -4),
          (935606702249, -10), (1987272877, 5)]
CGH_T2 = [(5, 0), (13, 0), (563, 0), (17475368544847, 154), (13561740531809, 120),
          (9461354987597, 94), (16556218163369, 2), (13707091918909, 143), (5609877309359, -6),
          (935606702249, -10), (14875476519749, -38), (1108967825921, 12), (15395725531427, 46),
          (2170161095393, 25), (1239053554603, -4), (16690620863071, 203), (4663421363459, 28),
          (2462223083147, -35), (7746014299613, 47), (17524177394617, 256), (11273815078217, 88),
          (10865903332033, 159), (7338481259891, -62), (16880979600449, 253)]

SMALL_RANGES = [(5, 20000), (100000, 120000), (1000000, 1020000), (1999000, 2000000),
                (86000000, 86010000)]


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1)
        f.write("\n")


def quot_hash(q: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(q, dtype="<u8").tobytes()).hexdigest()


def _two_squares_odd(p):
    """a with a^2 + b^2 = p, a odd, a = 1 (mod 4) (find.cl:46-65 normalisation)."""
    import math
    z = 2
    while pow(z, (p - 1) // 2, p) != p - 1:
        z += 1
    x = pow(z, (p - 1) // 4, p)
    r0, r1 = p, x
    while r1 * r1 > p:
        r0, r1 = r1, r0 % r1
    a, b = r1, math.isqrt(p - r1 * r1)
    assert a * a + b * b == p
    odd = a if a % 2 else b
    return -odd if odd % 4 == 3 else odd


def _eisenstein_c(p):
    """c with c^2 + 27 d^2 = 4p, c = 1 (mod 3) (find.cl:68-88 normalisation)."""
    import math
    g = 2
    while pow(g, (p - 1) // 3, p) == 1:
        g += 1
    w = pow(g, (p - 1) // 3, p)
    s = (2 * w + 1) % p
    if s % 2 == 0:
        s = p - s
    r0, r1 = 2 * p, s
    while r1 * r1 >= 4 * p:
        r0, r1 = r1, r0 % r1
    x, y = r1, math.isqrt((4 * p - r1 * r1) // 3)
    assert x * x + 3 * y * y == 4 * p
    for u, v in ((x, y), ((x - 3 * y) // 2, (x + y) // 2), ((x + 3 * y) // 2, (x - y) // 2)):
        if v % 3 == 0:
            c = abs(u)
            assert c * c + 27 * (v // 3) ** 2 == 4 * p
            return -c if c % 3 == 2 else c
    raise AssertionError(p)


def acu_values(p, t):
    """(acu1, acu2) for the oracle's finalize: (u, c) for type 0 (u = +-c with
    u = umod (mod 3), one of the valid u), (a, 0) for type 1."""
    if t == 1:
        return _two_squares_odd(p), 0
    if t == 0:
        c = _eisenstein_c(p)
        umod = 1 if ((p - 1) // 6) % 2 == 0 else 2
        u = abs(c) if abs(c) % 3 == umod else -abs(c)
        return u, c
    return 0, 0


def type_targets(ps) -> tuple[list[int], dict[int, int]]:
    """Types of the test primes of one range and each type's target, taken
    from the range's first prime of that type (cl_wilson.cpp:1552-1586)."""
    typ = [0 if int(p) % 3 == 1 else (1 if int(p) % 12 == 5 else 2) for p in ps]
    T = {}
    for p, t in zip(ps, typ):
        T.setdefault(t, (int(p) - 1) // (6, 4, 2)[t])
    return typ, T


MFMA_CASES = [  # (pmin, pmax, [window centre: an exponent-run boundary or a target])
    (20_000_000_000_000, 20_000_000_001_200, ["T2/2", "T2/3", "T1/2", "T0-"]),
    (24_876_000_000_000, 24_876_000_001_600, ["T2/2", "T0-"]),
    # 4-limb band (p above 2.49e13) on the matrix cores, up to its limit 2^47
    (30_000_000_000_000, 30_000_000_001_600, ["T2/2", "T0-"]),
    (100_000_000_000_000, 100_000_000_001_800, ["T2/3", "T1/2"]),
    (140_737_000_000_000, 140_737_000_001_800, ["T2/2", "T0-"]),
]
MFMA_HALF = 20_000_000  # windows of 4e7 integers (~1.4 M primes, ~10,800 sub-chunks)


def mfma_windows():
    out = []
    for plo, phi, centres in MFMA_CASES:
        ps = O.primes_range(plo, phi)
        typ, T = type_targets(ps)
        targets = np.array([T[t] for t in typ], dtype=np.uint64)
        for c in centres:
            if c.endswith("-"):  # the window ends at the type-0 target
                qhi = T[0] + 1
                qlo = qhi - 2 * MFMA_HALF
            else:
                t, d = int(c[1]), int(c[3])
                m = T[t] // d
                qlo, qhi = m - MFMA_HALF, m + MFMA_HALF
            t0 = time.time()
            vals = O.window_product(ps, targets, qlo, qhi)
            print("mfma window", plo, c, qlo, qhi, round(time.time() - t0, 1), flush=True)
            out.append(dict(pmin=plo, pmax=phi, at=c, qlo=qlo, qhi=qhi, p=[int(x) for x in ps],
                            type_targets={str(k): v for k, v in T.items()}, value=[str(v) for v in vals]))
    dump("mfma_windows.json", out)


def _cfg2_window(args):
    ps, targets, a, b = args
    return O.window_product(np.array(ps, dtype=np.uint64), np.array(targets, dtype=np.uint64), a, b, nthreads=1)


CFG2 = (100_000_000_000, 100_010_000_000)


def cfg2_samples(nsample=16, width=500_000_000, workers=8):
    from multiprocessing import Pool
    lo, hi = CFG2
    allp = O.primes_range(lo, hi)
    typ, T = type_targets(allp)
    rng = np.random.default_rng(2)
    idx = sorted(set([0, len(allp) - 1] + [int(i) for i in rng.choice(len(allp), nsample - 2, replace=False)]))
    ps = [int(allp[i]) for i in idx]
    ty = [typ[i] for i in idx]
    targets = [T[t] for t in ty]
    maxT = max(targets)
    jobs = [(ps, targets, a, min(a + width, maxT + 1)) for a in range(2, maxT + 1, width)]
    t0 = time.time()
    acc = [1] * len(ps)
    with Pool(workers) as pool:
        for k, vals in enumerate(pool.imap(_cfg2_window, jobs)):
            acc = [x * v % (p * p) for x, v, p in zip(acc, vals, ps)]
            if k % 10 == 0:
                print("cfg2 window", k, "of", len(jobs), round(time.time() - t0), "s", flush=True)
    own = np.array([(p - 1) // (6, 4, 2)[t] for p, t in zip(ps, ty)], dtype=np.uint64)
    itv = O.iterate(np.array(ps, dtype=np.uint64), np.array(targets, dtype=np.uint64), own, acc)
    rows = []
    for i, p, t, v in zip(idx, ps, ty, itv):
        acu1, acu2 = acu_values(p, t)
        q, r = O.finalize(p, t, v, acu1, acu2)
        assert r == 0, (p, r)
        rows.append(dict(index=i, p=p, type=t, quot=q, w=O.wilson_quotient(p, q)))
    dump("cfg2_oracle.json", dict(pmin=lo, pmax=hi, tpcount=len(allp),
                                  type_targets={str(k): v for k, v in T.items()}, samples=rows))
    print("cfg2 samples", round(time.time() - t0), "s", flush=True)


CFG4_RANGES = [0, 30, 60, 100, 140, 180, 200, 215, 221, 222, 240, 255]  # 222 on: the 4-limb band
CFG4_WINDOWS = [(1_000_000_000, 1_040_000_000), (2_000_000_000_000, 2_000_040_000_000)]


def cfg4_windows():
    sys.path.insert(0, os.path.join(HERE, "..", ".."))
    from clwilson_amd.workloads import batch_ranges
    ranges = batch_ranges()
    rows = []
    for r in CFG4_RANGES:
        lo, hi = ranges[r]
        allp = O.primes_range(lo, hi)
        typ, T = type_targets(allp)
        n = len(allp)
        idx = sorted(set([0, 1, 2, n // 3, n // 2, n - 2, n - 1]))
        ps = np.array([allp[i] for i in idx], dtype=np.uint64)
        targets = np.array([T[typ[i]] for i in idx], dtype=np.uint64)
        vals = []
        for qlo, qhi in CFG4_WINDOWS:
            t0 = time.time()
            vals.append([str(v) for v in O.window_product(ps, targets, qlo, qhi)])
            print("cfg4 range", r, qlo, round(time.time() - t0, 1), flush=True)
        rows.append(dict(range=r, pmin=lo, pmax=hi, tpcount=n, index=idx, p=[int(x) for x in ps],
                         targets=[int(x) for x in targets], windows=vals))
    dump("cfg4_windows.json", dict(windows=CFG4_WINDOWS, ranges=rows))


LARGE_Q = ((20_000_000_000_000, 20_000_000_100_000), (10**12, 10**12 + 2 * 10**9))


def large_q_window():
    """Oracle window products for sampled test primes of [2e13, 2e13+1e5)
    over the sieving primes of [1e12, 1e12+2e9), all above sqrt(T): the
    window of the batch's run-product carry test."""
    (lo, hi), (qlo, qhi) = LARGE_Q
    allp = O.primes_range(lo, hi)
    typ, T = type_targets(allp)
    n = len(allp)
    idx = sorted(set([0, 1, 2, n // 4, n // 2, 3 * n // 4, n - 2, n - 1]))
    ps = np.array([allp[i] for i in idx], dtype=np.uint64)
    targets = np.array([T[typ[i]] for i in idx], dtype=np.uint64)
    t0 = time.time()
    vals = O.window_product(ps, targets, qlo, qhi)
    print("large-q window", round(time.time() - t0, 1), flush=True)
    dump("large_q_window.json", dict(pmin=lo, pmax=hi, qlo=qlo, qhi=qhi, tpcount=n, index=idx,
                                     p=[int(x) for x in ps], targets=[int(x) for x in targets],
                                     value=[str(v) for v in vals]))


def main():
    if sys.argv[1:] == ["--large-q"]:
        return large_q_window()
    if sys.argv[1:] == ["--direct"]:
        return direct_only()
    if sys.argv[1:] == ["--mfma"]:
        return mfma_windows()
    if sys.argv[1:] == ["--cfg2"]:
        return cfg2_samples()
    if sys.argv[1:] == ["--cfg4"]:
        return cfg4_windows()
    dump("kat_selftest.json", [dict(pmin=a, pmax=b, results=c, checksum=f"{d:016X}", pi_maxt=e, lastp=f,
                                    lastw=g) for a, b, c, d, e, f, g in KATS])
    dump("cgh_tables.json", dict(table1=CGH_T1, table2=CGH_T2))
    small = []
    for lo, hi in SMALL_RANGES:
        t = time.time()
        arr, st = O.wilson_range(lo, hi)
        assert st["rc"] == 0
        small.append(dict(pmin=lo, pmax=hi, checksum=f"{st['checksum']:016X}", pi_maxt=st["totalcount"],
                          results=st["resultcount"], p=[int(x) for x in arr["p"]],
                          quot=[int(x) for x in arr["quot"]]))
        print("small", lo, hi, round(time.time() - t, 2), flush=True)
    dump("oracle_small.json", small)
    # Window products at config-2 / config-3 magnitudes: a few test primes of
    # each type, sieving primes in a fixed window of the target axis.
    wins = []
    for (plo, phi, qlo, qhi) in [(100000000000, 100000000400, 16666000000, 16666400000),
                                 (20000000000000, 20000000000400, 3333333000000, 3333333300000),
                                 (20000000000000, 20000000000400, 9999999700000, 10000000000056),
                                 (1000000000000000, 1000000000000300, 166666666500000, 166666666700000),
                                 (1000000000000000000, 1000000000000000200, 166666666666600000,
                                  166666666666700000)]:
        ps = O.primes_range(plo, phi)
        T = {}
        for p in ps:
            p = int(p)
            t = 0 if p % 3 == 1 else (1 if p % 12 == 5 else 2)
            T.setdefault(t, (p - 1) // (6, 4, 2)[t])
        typ = [0 if int(p) % 3 == 1 else (1 if int(p) % 12 == 5 else 2) for p in ps]
        targets = np.array([T[t] for t in typ], dtype=np.uint64)
        t0 = time.time()
        vals = O.window_product(ps, targets, qlo, qhi)
        print("window", plo, qlo, qhi, round(time.time() - t0, 2), flush=True)
        own = np.array([(int(p) - 1) // (6, 4, 2)[t] for p, t in zip(ps, typ)], dtype=np.uint64)
        itv = O.iterate(ps, targets, own, vals)
        fin = []
        for p, t, v in zip(ps, typ, itv):
            acu1, acu2 = acu_values(int(p), t)
            fin.append(list(O.finalize(int(p), t, v, acu1, acu2)))
        wins.append(dict(pmin=plo, pmax=phi, qlo=qlo, qhi=qhi, p=[int(x) for x in ps],
                         type_targets={str(k): v for k, v in T.items()}, value=[str(v) for v in vals],
                         iterated=[str(v) for v in itv], finalized=fin))
    dump("windows.json", wins)
    if "--cfg1" in sys.argv:
        t = time.time()
        arr, st = O.wilson_range(1000000, 2000000)
        assert st["rc"] == 0
        res = [dict(p=int(p), w=O.wilson_quotient(int(p), int(q))) for p, q in zip(arr["p"], arr["quot"])
               if int(q) == 0 or min(int(q), int(p) - int(q)) < 1000]
        dump("cfg1.json", dict(pmin=1000000, pmax=2000000, tpcount=len(arr["p"]),
                               checksum=f"{st['checksum']:016X}", pi_maxt=st["totalcount"],
                               type_targets=st["typeTarget"], tpcnt=st["tpcnt"],
                               resultcount=st["resultcount"], results=res,
                               quot_sha256=quot_hash(arr["quot"]), p_sha256=quot_hash(arr["p"])))
        print("cfg1", round(time.time() - t, 1), hex(st["checksum"]), flush=True)
    if "--direct" in sys.argv:
        direct_only()


def direct_only():
    out = []
    for lo, hi in [(5, 1000000), (1000000, 2000000)]:
        t = time.time()
        ps = O.primes_range(lo, hi)
        q, r = O.direct_quotients(ps)
        assert not r.any()
        checksum = int((ps.astype(object) + q.astype(object) + r.astype(object)).sum()) % 2**64
        out.append(dict(pmin=lo, pmax=hi, tpcount=len(ps), checksum=f"{checksum:016X}", quot_sha256=quot_hash(q),
                        samples=[[int(ps[i]), int(q[i])] for i in range(0, len(ps), 500)]))
        print("direct", lo, hi, round(time.time() - t, 1), flush=True)
    dump("direct_lt_2e6.json", out)


if __name__ == "__main__":
    main()
