"""Writes the golden fixtures of tests/golden/ with the CPU oracle.

    python tests/golden/make_golden.py [--cfg1] [--direct]

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


def main():
    if sys.argv[1:] == ["--direct"]:
        return direct_only()
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
