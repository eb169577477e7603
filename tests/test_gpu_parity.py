"""GPU parity: the HIP path through the C ABI against the oracle's fixtures,
the reference's self-test answers and the CGH published Wilson quotients.

Integer work, so every comparison is bit-exact.
"""
import hashlib
import sys

import numpy as np
import pytest

import clwilson_amd as cw

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _kat_check(res, kat):
    assert res.resultcount == kat["results"]
    assert res.checksum_hex == kat["checksum"]
    assert res.totalcount == kat["pi_maxt"]
    assert (res.lastp, res.lastw) == (kat["lastp"], kat["lastw"])


def test_selftest_case5_small_iterations(dev, golden):
    """cl_wilson.cpp:2076-2089: [200, 564), Wilson prime 563."""
    kat = golden("kat_selftest.json")[4]
    res = dev.wilson_range(kat["pmin"], kat["pmax"])
    _kat_check(res, kat)
    assert res.results[-1] == "563 is a Wilson prime\n"


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_small_ranges_bit_exact(dev, golden, mode):
    """Every prime's Wilson quotient equals the oracle's, for the lane-per-prime
    and wave-per-prime kernels and the reference's one-prime-per-launch path."""
    for case in golden("oracle_small.json"):
        say("range", case["pmin"], case["pmax"], "mode", mode)
        res = dev.wilson_range(case["pmin"], case["pmax"], mode=mode)
        assert [int(x) for x in res.p] == case["p"]
        assert [int(x) for x in res.quot] == case["quot"]
        assert not res.rem.any()
        assert res.checksum_hex == case["checksum"]
        assert res.totalcount == case["pi_maxt"]


@pytest.mark.parametrize("mode", [0, 1])
def test_config1_full_range(dev, golden, mode):
    """[1e6, 2e6): 70,435 primes; the quotient vector hashes to the oracle's,
    through the automatic mode (a one-range batch at this size) and the
    session's lane-per-test-prime kernel."""
    g = golden("cfg1.json")
    res = dev.wilson_range(g["pmin"], g["pmax"], mode=mode)
    assert len(res.p) == g["tpcount"]
    assert hashlib.sha256(np.ascontiguousarray(res.p, dtype="<u8").tobytes()).hexdigest() == g["p_sha256"]
    assert hashlib.sha256(np.ascontiguousarray(res.quot, dtype="<u8").tobytes()).hexdigest() == g["quot_sha256"]
    assert res.checksum_hex == g["checksum"]
    assert res.totalcount == g["pi_maxt"]
    assert res.type_targets == g["type_targets"] and res.tpcnt == g["tpcnt"]
    got = [(int(p), cw.wilson_quotient(int(p), int(q))) for p, q in zip(res.p, res.quot)
           if cw.format_result(int(p), int(q))]
    assert got == [(r["p"], r["w"]) for r in g["results"]]


def test_selftest_case6_large_iterations(dev, golden):
    """cl_wilson.cpp:2093-2106: [86000000, 87467200), near-Wilson 87467099 -2."""
    kat = golden("kat_selftest.json")[5]
    _kat_check(dev.wilson_range(kat["pmin"], kat["pmax"]), kat)


def _cgh_cases(golden_loader, limit_lo, limit_hi):
    g = golden_loader("cgh_tables.json")
    seen = {}
    for p, w in g["table1"] + g["table2"]:
        seen[p] = w
    return [(p, w) for p, w in sorted(seen.items()) if limit_lo <= p < limit_hi]


def test_cgh_values_up_to_1e11(dev, golden):
    for p, w in _cgh_cases(golden, 0, 10**11):
        res = dev.wilson_range(p, p + 1) if p >= 5 else None
        if res is None:
            continue
        say("cgh", p, w, res.lastw, round(res.seconds, 3))
        assert res.resultcount == 1 and res.lastw == w, p


@pytest.mark.parametrize("mode", [1, 2])
def test_window_products_all_limb_bands(dev, golden, mode):
    """Canonical products over a window of sieving primes at p ~ 1e11, 2e13,
    1e15, 1e18 (3, 3, 4, 5 limbs), then iterate and the final check."""
    for w in golden("windows.json"):
        say("window", w["pmin"], w["qlo"], w["qhi"], "mode", mode)
        with cw.Session(dev, w["pmin"], w["pmax"], mode=mode) as s:
            assert [int(x) for x in s.primes()] == w["p"]
            s.run(w["qlo"], w["qhi"])
            assert s.residues() == [int(v) for v in w["value"]]
            s.iterate()
            assert s.residues() == [int(v) for v in w["iterated"]]
            # a window is not T_p!, so rem != 0 is expected here
            res = s.finalize(allow_check_failure=True)
            assert [[int(q), int(r)] for q, r in zip(res.quot, res.rem)] == w["finalized"]


def test_window_split_invariance(dev):
    """Size-independent property: the product over [2, maxT] does not depend
    on how the sieving axis is cut into windows and segments."""
    lo, hi = 10**9, 10**9 + 20000
    ref = dev.wilson_range(lo, hi)
    for cuts, span in [((2, 12345, 4_000_000, 10**8), 0), ((2, 61441 * 3, 166666667), 61440 * 5)]:
        with cw.Session(dev, lo, hi, segment_span=span) as s:
            edges = list(cuts) + [s.maxtarget + 1]
            for a, b in zip(edges[:-1], edges[1:]):
                s.run(a, b)
            s.iterate()
            res = s.finalize()
        assert np.array_equal(res.quot, ref.quot) and res.checksum == ref.checksum
        assert res.totalcount == ref.totalcount


def test_chunking_invariance(dev):
    """Different wave targets (chunk counts K) give identical residues."""
    lo, hi = 3 * 10**10, 3 * 10**10 + 40000
    outs = []
    for waves in (64, 4096, 100000):
        with cw.Session(dev, lo, hi, target_waves=waves) as s:
            s.run(2, 10**8)
            outs.append(s.residues())
    assert outs[0] == outs[1] == outs[2]


def test_edge_ranges(dev):
    """Smallest ranges: type targets of 1 (p = 5, 7), only side records
    (p = 11, T = 5), and one prime of each type."""
    import oracle as O
    for lo, hi in [(5, 6), (5, 8), (7, 8), (11, 12), (5, 14), (13, 14), (5, 100), (101, 103)]:
        res = dev.wilson_range(lo, hi)
        arr, st = O.wilson_range(lo, hi)
        assert [int(x) for x in res.quot] == [int(x) for x in arr["quot"]], (lo, hi)
        assert res.checksum == st["checksum"] and res.totalcount == st["totalcount"]


@pytest.mark.slow
def test_cgh_values_large(dev, golden):
    for p, w in _cgh_cases(golden, 10**11, 10**12):
        res = dev.wilson_range(p, p + 1)
        say("cgh", p, w, res.lastw, round(res.seconds, 2))
        assert res.resultcount == 1 and res.lastw == w, p


@pytest.mark.slow
@pytest.mark.parametrize("k", [0, 1, 2, 3, 6])
def test_selftest_large_primes(dev, golden, k):
    """cl_wilson.cpp:2007-2072, 2110-2123: single primes 1.1e12-1.75e13 (all
    three types; case 7 is in config 3's magnitude band)."""
    kat = golden("kat_selftest.json")[k]
    res = dev.wilson_range(kat["pmin"], kat["pmax"])
    say("kat", k + 1, round(res.seconds, 2))
    _kat_check(res, kat)


@pytest.mark.parametrize("lo,hi", [(2, 100), (5, 20000), (7, 8), (61439, 61441), (1_000_000, 3_000_000),
                                   (2**32 - 5000, 2**32 + 5000), (10**14, 10**14 + 200_000),
                                   (2**62 - 20_000, 2**62), (759_250_000, 759_260_000)])
def test_device_primes_equal_host_primes(dev, lo, hi):
    """The device-sieved test-prime list equals the host segmented sieve."""
    assert np.array_equal(dev.primes(lo, hi), cw.host_primes(lo, hi))


@pytest.mark.parametrize("engine", ["session", "batch"])
def test_every_prime_below_2e6_vs_definition(dev, golden, engine):
    """Every prime p < 2e6 against (p-1)! = 2*3*...*(p-1) mod p^2 computed
    by definition (direct_lt_2e6.json): quotient-vector hash and checksum of
    [5, 1e6) and [1e6, 2e6), through the session path and the batched engine."""
    for d in golden("direct_lt_2e6.json"):
        if engine == "session":
            r = dev.wilson_range(d["pmin"], d["pmax"], mode=1)
        else:
            with cw.Batch(dev, [(d["pmin"], d["pmax"])]) as b:
                r = b.search()[0]
        assert len(r.p) == d["tpcount"] and r.checksum_hex == d["checksum"]
        assert hashlib.sha256(np.ascontiguousarray(r.quot, dtype="<u8").tobytes()).hexdigest() == d["quot_sha256"]
