"""GPU parity on the benchmark workloads and on the matrix-core run products.

* the matrix cores (k_mfma_runs, byte tables from HBM and built per
  workgroup) against the oracle's window products in the 3-limb band (p ~ 2e13
  and its top) and the 4-limb band (p ~ 3e13, 1e14, just below 2^47), on
  windows long enough for the engine to pick them;
* every prime in [1e11, 1e11+1e7) (config 2): the full search through the
  batched engine, sampled primes against the oracle's full plain method;
* the 256-range batch (config 4): two windows of sieving primes, sampled
  test primes of ranges in the 3- and 4-limb bands against the oracle;
* the 2,000-request serving stream (config 5): a bounded replay, every
  completed request equal to its single-range run, the HBM budget held.

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


@pytest.mark.parametrize("mode", [0, 2])
def test_matrix_core_windows_vs_oracle(dev, golden, mode):
    """Window products through a batch whose run products go to the matrix
    cores (mode 0: byte tables in HBM; mode 2: built per workgroup in LDS)
    equal the oracle's (mullarge.cl:36-64 restated): 3 limbs at p ~ 2e13
    (lazy run accumulator) and p ~ 2.4876e13 (conditional subtraction kept),
    4 limbs at p ~ 3e13, 1e14 and just below 2^47, on windows across T/2 and
    T/3 exponent-run boundaries and type targets."""
    for w in golden("mfma_windows.json"):
        say("window", w["pmin"], w["at"], w["qlo"], w["qhi"], "mode", mode)
        with cw.Batch(dev, [(w["pmin"], w["pmax"])], run_products=mode) as b:
            assert b.info[0].limbs == (3 if w["pmax"] < 24_880_000_000_000 else 4)
            assert b.info[0].tpcount >= 32
            assert list(b.info[0].typeTarget) == [w["type_targets"][str(t)] for t in range(3)]
            b.run(w["qlo"], w["qhi"])
            ks = b.kernel_stats()
            assert ks["mfma_int8_macs"] > 0, "the window did not reach the matrix cores"
            assert b.residues(0) == [int(v) for v in w["value"]]


def test_matrix_cores_stop_at_2_47(dev):
    """Test primes at or above 2^47 keep the scalar run products (their
    sub-chunk products no longer fit the byte tables): a window there runs
    no matrix-core work and equals the same window with the matrix cores
    switched off, while a range just below 2^47 takes them."""
    q = (35_184_230_000_016, 35_184_270_000_016)
    for lo, on in ((140_737_488_355_400, False), (140_737_000_000_000, True)):
        out = []
        for mode in (0, 1):
            with cw.Batch(dev, [(lo, lo + 1800)], run_products=mode) as b:
                assert b.info[0].limbs == 4
                b.run(*q)
                out.append(b.residues(0))
                macs = b.kernel_stats()["mfma_int8_macs"]
                say("p", lo, "mode", mode, "int8 MACs", macs)
                assert (macs > 0) == (on and mode == 0)
        assert out[0] == out[1]


@pytest.mark.slow
def test_config2_full_range(dev, golden):
    """[1e11, 1e11+1e7), 394,401 test primes, full search through the batched
    engine (matrix-core run products): 16 sampled primes equal the oracle's
    full plain method (cfg2_oracle.json); checksum, quotient hash and
    pi(maxtarget) = pi(5e10) as recorded (cfg2_gpu.json)."""
    g = golden("cfg2_gpu.json")
    o = golden("cfg2_oracle.json")
    with cw.Batch(dev, [(g["pmin"], g["pmax"])]) as b:
        M = b.maxtarget
        for a in range(2, M + 1, 5_000_000_000):
            b.run(a, min(a + 5_000_000_000, M + 1))
            say("covered", b.covered())
        applied = b.applied()
        macs = b.kernel_stats()["mfma_int8_macs"]
        b.iterate()
        b.finalize()
        r = b.results(0)
    assert macs > 0
    assert applied == g["pi_maxt"] == 2_119_654_578  # pi(5e10 + 9), every prime once
    assert len(r.p) == g["tpcount"] == o["tpcount"]
    for s in o["samples"]:
        assert int(r.p[s["index"]]) == s["p"]
        assert int(r.quot[s["index"]]) == s["quot"], s
    assert r.checksum_hex == g["checksum"] and r.results == g["results"]
    assert hashlib.sha256(np.ascontiguousarray(r.quot, dtype="<u8").tobytes()).hexdigest() == g["quot_sha256"]


def test_config4_batch_windows_vs_oracle(dev, golden):
    """All 256 ranges of the batched workload in one batch with segments of
    1e14 factor applications (so the 3-limb band's run products go to the
    matrix cores, tables built per workgroup); after each of two windows of
    sieving primes, sampled test primes of 12 ranges (3- and 4-limb bands)
    hold the oracle's products over the windows so far."""
    from clwilson_amd.workloads import batch_ranges
    g = golden("cfg4_windows.json")
    ranges = batch_ranges()
    with cw.Batch(dev, ranges, segment_apps=10**14) as b:
        say("config-4 batch:", sum(st.tpcount for st in b.info), "test primes")
        bands = {b.info[c["range"]].limbs for c in g["ranges"]}
        assert bands == {3, 4}
        acc = {c["range"]: [1] * len(c["p"]) for c in g["ranges"]}
        for k, (qlo, qhi) in enumerate(g["windows"]):
            b.run(qlo, qhi)
            ks = b.kernel_stats()
            say("window", qlo, qhi, "int8 MACs", ks["mfma_int8_macs"], "prime apps", ks["mul_prime_apps"])
            assert ks["mfma_int8_macs"] > 0
            for c in g["ranges"]:
                r = c["range"]
                assert b.info[r].tpcount == c["tpcount"]
                res = b.residues(r)
                acc[r] = [x * int(v) % (p * p) for x, v, p in zip(acc[r], c["windows"][k], c["p"])]
                assert [res[i] for i in c["index"]] == acc[r], (r, k)


def test_config5_serving_replay(dev):
    """The 2,000-request stream (workloads.serving_requests(), seed 99)
    replayed in real time for 20 s under a 64 GiB budget: every request
    completed in that window equals the same range run alone, and the
    budget holds."""
    from clwilson_amd.serve import Server
    from clwilson_amd.workloads import serving_requests
    reqs = serving_requests()
    budget = 64 << 30
    srv = Server(hbm_budget=budget)
    try:
        rep = srv.serve(reqs, duration=20.0)
    finally:
        srv.close()
    say({k: rep[k] for k in ("arrived", "admitted", "completed", "p50_latency_s", "p99_latency_s", "hbm_peak")})
    assert rep["completed"] >= 5 and rep["hbm_peak"] <= budget
    byid = {r.rid: r for r in reqs}
    for c in srv.completed:
        r = byid[c.rid]
        a = dev.wilson_range(r.pmin, r.pmax)
        assert np.array_equal(c.result.p, a.p) and np.array_equal(c.result.quot, a.quot), r
        assert c.result.checksum == a.checksum and c.result.results == a.results


# Costa-Gerbicz-Harvey Table 2 values and self-test cases 4 and 7 in config 3's
# magnitude band (1.0e13-1.75e13): their targets are at most 4.22e12 (types 0
# and 1), so one sieve pass serves all of them.
CGH_BAND = [10865903332033, 11273815078217, 13561740531809, 13707091918909, 14875476519749, 16556218163369,
            16690620863071, 16880979600449, 17475368544847, 17524177394617]


@pytest.mark.slow
@pytest.mark.parametrize("mode", [0, 2])
def test_cgh_band_through_matrix_cores(dev, golden, mode):
    """Full searches through one batch whose 3-limb run products go to the
    matrix cores (mode 0: byte tables in HBM, mode 2: built per workgroup):
    every published w_p of the ten primes of config 3's band (CGH Table 2,
    self-test cases 4 and 7, cl_wilson.cpp:2059-2072, 2110-2123). Mode 0 runs
    self-test case 7 as its whole range [17524177394450, 17524177394618)
    (targets up to 8.76e12), so its checksum, result count and pi(maxT) are
    checked too; mode 2 runs the ten primes alone."""
    g = golden("cgh_tables.json")
    want = {p: w for p, w in g["table1"] + g["table2"]}
    kat7 = golden("kat_selftest.json")[6]
    ranges = [(p, p + 1) for p in CGH_BAND[:-1]]
    ranges.append((kat7["pmin"], kat7["pmax"]) if mode == 0 else (CGH_BAND[-1], CGH_BAND[-1] + 1))
    with cw.Batch(dev, ranges, run_products=mode) as b:
        assert all(st.limbs == 3 for st in b.info)
        M = b.maxtarget
        step = 500_000_000_000
        for a in range(2, M + 1, step):
            b.run(a, min(a + step, M + 1))
            say("mode", mode, "covered", b.covered(), "of", M + 1)
        macs = b.kernel_stats()["mfma_int8_macs"]
        applied = b.applied()  # every prime up to the largest target, once
        b.iterate()
        b.finalize()
        res = [b.results(r) for r in range(len(ranges))]
    assert macs > 0, "no run product reached the matrix cores"
    if mode == 0:  # case 7 holds the batch's largest target
        assert applied == kat7["pi_maxt"]
    for p, r in zip(CGH_BAND, res):
        assert r.lastp == p and r.lastw == want[p], (p, r.lastw, want[p])
    if mode == 0:
        r7 = res[-1]
        assert r7.resultcount == kat7["results"] and r7.checksum_hex == kat7["checksum"]
        assert (r7.lastp, r7.lastw) == (kat7["lastp"], kat7["lastw"])
