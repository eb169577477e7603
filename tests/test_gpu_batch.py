"""GPU parity of the batched engine: many ranges over one shared sieve.

Every range of a batch must give exactly what a single-range run gives (the
session path, itself pinned to the oracle and the reference's self-test
answers), whatever the other ranges in the batch, the window splits, the
segment span, the graph replay or the work-queue chunking. Integer work, so
every comparison is bit-exact.
"""
import sys

import numpy as np
import pytest

import clwilson_amd as cw

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def _same(a, b):
    assert (a.pmin, a.pmax) == (b.pmin, b.pmax)
    assert np.array_equal(a.p, b.p)
    assert np.array_equal(a.quot, b.quot)
    assert np.array_equal(a.rem, b.rem)
    assert a.checksum == b.checksum
    assert (a.resultcount, a.lastp, a.lastw) == (b.resultcount, b.lastp, b.lastw)
    assert a.results == b.results
    assert a.type_targets == b.type_targets and a.tpcnt == b.tpcnt


# Ranges of every magnitude up to ~1e9 (targets up to 5e8: full searches
# take a second): the 2-limb band and, above 759,250,125, the 3-limb band.
MIX = [(5, 20000), (100000, 120000), (1000000, 1020000), (1999000, 2000000), (86000000, 86010000),
       (759245000, 759255000), (1000000000, 1000004000), (999999000, 1000000000), (563, 564)]


def test_batch_matches_oracle_fixtures(dev, golden):
    """A batch of the oracle's small ranges returns the oracle's quotients."""
    cases = golden("oracle_small.json")
    say("batch of", len(cases), "oracle ranges")
    with cw.Batch(dev, [(c["pmin"], c["pmax"]) for c in cases]) as b:
        res = b.search()
    for c, r in zip(cases, res):
        assert [int(x) for x in r.p] == c["p"]
        assert [int(x) for x in r.quot] == c["quot"]
        assert not r.rem.any()
        assert r.checksum_hex == c["checksum"]


def test_batch_selftest_case5(dev, golden):
    """cl_wilson.cpp:2076-2089 inside a batch: [200, 564), Wilson prime 563."""
    kat = golden("kat_selftest.json")[4]
    with cw.Batch(dev, [(kat["pmin"], kat["pmax"]), (100000, 100100)]) as b:
        r = b.search()[0]
    assert r.resultcount == kat["results"]
    assert r.checksum_hex == kat["checksum"]
    assert (r.lastp, r.lastw) == (kat["lastp"], kat["lastw"])


@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_single_ranges(dev, graph):
    """Mixed magnitudes and both limb bands: each range equals its own run."""
    say("mixed batch, graph", graph)
    with cw.Batch(dev, MIX, graph=graph) as b:
        got = b.search()
    for (lo, hi), r in zip(MIX, got):
        say("  range", lo, hi)
        _same(r, dev.wilson_range(lo, hi))


def test_batch_alone_vs_mix(dev):
    """A range's results do not depend on the batch around it."""
    with cw.Batch(dev, MIX) as b:
        mixed = b.search()
    for k in (0, 4, 6):
        with cw.Batch(dev, [MIX[k]]) as b:
            _same(b.search()[0], mixed[k])


def test_batch_windows_and_span_invariance(dev):
    """Window splits (including ones that cut 2, 3, 5 off) and a short
    segment span (many graph replays per window) change nothing."""
    ranges = MIX[2:7]
    with cw.Batch(dev, ranges) as b:
        ref = b.search()
    cuts = [2, 4, 6, 1000, 65537, 12345678, 1 << 62]
    say("windows", cuts)
    with cw.Batch(dev, ranges, segment_span=3 * 61440) as b:
        for a, z in zip(cuts, cuts[1:]):
            b.run(a, z)
        b.iterate()
        b.finalize()
        got = [b.results(r) for r in range(len(ranges))]
    for x, y in zip(ref, got):
        _same(x, y)


def test_batch_many_small_ranges(dev):
    """256 narrow ranges (768 groups, several chunks per group)."""
    rng = np.random.default_rng(7)
    starts = sorted(int(x) for x in rng.integers(10**6, 6 * 10**8, size=256))
    ranges = [(s, s + int(w)) for s, w in zip(starts, rng.integers(200, 3000, size=256))]
    say("256 ranges")
    with cw.Batch(dev, ranges) as b:
        got = b.search()
    for k in (0, 1, 100, 200, 255):
        _same(got[k], dev.wilson_range(*ranges[k]))


@pytest.mark.parametrize("window", [(2, 300000000), (150000000, 2000000000)])
def test_batch_window_residues_all_bands(dev, window):
    """Partial products over a window of sieving primes for ranges in the 2,
    3 and 4-limb bands equal the session's over the same window."""
    ranges = [(600000000, 600002000), (3 * 10**12, 3 * 10**12 + 2000), (10**14, 10**14 + 3000)]
    qlo, qhi = window
    with cw.Batch(dev, ranges) as b:
        assert [st.limbs for st in b.info] == [2, 3, 4]
        b.run(qlo, qhi)
        for r, (lo, hi) in enumerate(ranges):
            say("window", window, "range", lo, hi)
            with cw.Session(dev, lo, hi) as s:
                s.run(qlo, qhi)
                assert b.residues(r) == s.residues()


def test_batch_run_products_carry_large_q(dev, golden):
    """Above sqrt(T) almost every sieving prime enters through a run product
    (one full Montgomery product per 128-prime sub-chunk, the digits reduced
    mod p^2), and the window residues equal the session path's, which applies
    every record as a product, and, for sampled test primes, the oracle's
    window products (mullarge.cl:36-64 restated)."""
    g = golden("large_q_window.json")
    lo, hi = g["pmin"], g["pmax"]
    qlo, qhi = g["qlo"], g["qhi"]
    with cw.Batch(dev, [(lo, hi)]) as b, cw.Session(dev, lo, hi) as s:
        b.run(qlo, qhi)
        s.run(qlo, qhi)
        ks = b.kernel_stats()
        res = b.residues(0)
        assert b.info[0].tpcount == g["tpcount"]
        assert [res[i] for i in g["index"]] == [int(v) for v in g["value"]]
        assert res == s.residues()
    say("prime apps", ks["mul_prime_apps"], "full products", ks["mul_record_apps"],
        "partial products", ks["mul_partial_products"])
    assert ks["mul_prime_apps"] > 0
    assert ks["mul_record_apps"] * 100 < ks["mul_prime_apps"]
    assert ks["mul_partial_products"] < 7 * ks["mul_prime_apps"]  # ~5.9 per prime at this size


def test_run_products_matrix_cores_equal_vector_units(dev):
    """The 3-limb band's run products as an int8 GEMM on the matrix cores
    (k_mfma_runs, the default, with the byte tables from HBM; and mode 2, the
    tables built by each workgroup in LDS) give the residues of the
    vector-unit reduction (k_batch_mul, mode 1) bit for bit: windows across
    exponent-run boundaries (T/2 of 1e12's type-2 target 5e11, and type 1's
    target 2.5e11), T itself and the large-q region of 2e13."""
    ranges = [(10**12, 10**12 + 20000), (3 * 10**12, 3 * 10**12 + 20000), (2 * 10**13, 2 * 10**13 + 20000)]
    # (1e6, 2e6) is too short for the matrix cores and is the first window
    # replayed from a captured graph; the next one allocates the byte tables,
    # so the step is captured again with them
    windows = [(2, 10**6), (10**6, 2 * 10**6), (2 * 10**11 + 5 * 10**10 - 10**9, 2 * 10**11 + 5 * 10**10 + 10**9),
               (5 * 10**11 - 10**9, 5 * 10**11 + 10**9), (10**12, 10**12 + 2 * 10**9)]
    got = {}
    for mode in (0, 1, 2):
        with cw.Batch(dev, ranges, run_products=mode) as b:
            for qlo, qhi in windows:
                b.run(qlo, qhi)
            got[mode] = ([b.residues(r) for r in range(len(ranges))], b.kernel_stats())
    for mode in (0, 2):
        for r in range(len(ranges)):
            assert got[mode][0][r] == got[1][0][r], f"mode {mode} range {ranges[r]}"
        say("mode", mode, "int8 MACs", got[mode][1]["mfma_int8_macs"], "prime apps", got[mode][1]["mul_prime_apps"])
        assert got[mode][1]["mfma_int8_macs"] > 0
        assert got[mode][1]["mul_prime_apps"] == got[1][1]["mul_prime_apps"]
    assert got[1][1]["mfma_int8_macs"] == 0


def test_run_products_matrix_cores_top_of_band(dev):
    """Test primes just below the 3-limb band's top (2 p^2 close to 2^90):
    k_mfma_runs keeps mul_full's conditional subtraction there (the lazy
    product needs N < 0.4998 R) and still equals the vector reduction."""
    ranges = [(24_876_000_000_000, 24_876_000_020_000)]
    qlo, qhi = 10**12, 10**12 + 2 * 10**9
    got = {}
    for mode in (0, 1):
        with cw.Batch(dev, ranges, run_products=mode) as b:
            assert b.info[0].limbs == 3 and 2 * (ranges[0][1] - 1) ** 2 > 0.9996 * 2**90
            b.run(qlo, qhi)
            got[mode] = (b.residues(0), b.kernel_stats())
    assert got[0][0] == got[1][0]
    assert got[0][1]["mfma_int8_macs"] > 0 and got[1][1]["mfma_int8_macs"] == 0


def test_batch_rejects_bad_ranges(dev):
    with pytest.raises(cw.ClwError):
        cw.Batch(dev, [(10, 10)])
    with pytest.raises(cw.ClwError):
        cw.Batch(dev, [(10**18, 10**18 + 100)])  # 5-limb band: sessions only
    with pytest.raises(cw.ClwError):
        cw.Batch(dev, [(100, 100 + 10**7 + 1)])


def test_autotune_keeps_results(dev, tmp_path):
    """The autotuner's trial segments leave every residue untouched and its
    choice (whatever it is) gives the untuned results bit for bit; a second
    tuning of the same shape comes from the cache."""
    ranges = MIX[1:7]
    with cw.Batch(dev, ranges) as b:
        ref = b.search()
    cache = str(tmp_path / "tune.txt")
    with cw.Batch(dev, ranges) as b:
        b.run(2, 1000)
        before = [b.residues(r) for r in range(len(ranges))]
        rep = b.autotune(5_000_000, cache)
        say("autotune", rep)
        assert rep and all(t["trials"] >= 3 and not t["from_cache"] for t in rep)
        assert all(t["per_lane"] in (1, 2, 3, 4, 5, 6) for t in rep)
        assert all(rate > 0 for t in rep for _, _, rate in t["candidates"])  # every trial did its work
        assert all(ms > 0 for t in rep for ms in t["candidate_ms"])
        assert [b.residues(r) for r in range(len(ranges))] == before
        b.run(1000, b.maxtarget + 1)
        b.iterate()
        b.finalize()
        for r in range(len(ranges)):
            _same(b.results(r), ref[r])
    with cw.Batch(dev, ranges) as b:
        rep2 = b.autotune(5_000_000, cache)
        assert all(t["from_cache"] for t in rep2)
        assert [(t["per_lane"], t["items_per_block"]) for t in rep2] == \
            [(t["per_lane"], t["items_per_block"]) for t in rep]


def test_autotune_matrix_cores_keeps_results(dev, tmp_path):
    """The matrix-core part of the autotuner (row slices of k_mfma_runs and
    the switch between the matrix cores and the vector units, both timed on
    trial segments at p ~ 2e13): its trials leave the residues untouched, a
    window after tuning equals the same window untuned bit for bit, and a
    second tuning of the same shape comes from the cache."""
    rng = [(2 * 10**13, 2 * 10**13 + 1_000_000)]
    w = (10**12, 10**12 + 2 * 10**9)
    with cw.Batch(dev, rng) as b:
        b.run(2, 1000)
        b.run(*w)
        ref = b.residues(0)
    cache = str(tmp_path / "tune.txt")
    with cw.Batch(dev, rng) as b:
        b.run(2, 1000)
        before = b.residues(0)
        rep = b.autotune(w[0], cache)
        say("autotune", rep)
        t = [x for x in rep if x["limbs"] == 3][0]
        assert t["mfma_slices"] in (8, 16) and t["mfma_min_rows"] >= 64
        cand = t["mfma_candidates"]
        assert len(cand) >= 4 and {s for s, _, _ in cand} >= {0, 8} and all(ms > 0 for _, _, ms in cand)
        # per-lane trials time the vector kernel alone on what the matrix cores leave it
        assert t["candidates"] and all(rate >= 0 for _, _, rate in t["candidates"])
        assert len(t["candidate_ms"]) == len(t["candidates"]) and all(ms > 0 for ms in t["candidate_ms"])
        assert b.residues(0) == before
        b.run(*w)
        assert b.residues(0) == ref
    with cw.Batch(dev, rng) as b:
        rep2 = b.autotune(w[0], cache)
        t2 = [x for x in rep2 if x["limbs"] == 3][0]
        assert t2["from_cache"]
        assert (t2["per_lane"], t2["items_per_block"], t2["mfma_slices"], t2["mfma_min_rows"]) == \
            (t["per_lane"], t["items_per_block"], t["mfma_slices"], t["mfma_min_rows"])


@pytest.mark.parametrize("per_lane", [1, 2, 3, 4, 5, 6])
def test_batch_every_per_lane_variant(dev, per_lane):
    """Every instantiation the autotuner may pick (test primes per lane)
    gives the same results: 2 and 3 limbs by full searches (5-6 per lane: 3
    limbs only), 4 limbs (1-5 per lane) by window residues against the
    session path."""
    ranges = MIX[5:8] if per_lane > 4 else MIX[3:8]  # 5-6 per lane: 3 limbs only
    say("per_lane", per_lane)
    with cw.Batch(dev, ranges, per_lane=per_lane) as b:
        got = b.search()
    for (lo, hi), r in zip(ranges, got):
        _same(r, dev.wilson_range(lo, hi))
    if per_lane <= 5:  # 4-limb variants: 1-5 per lane
        lo4, hi4 = 10**14, 10**14 + 3000
        with cw.Batch(dev, [(lo4, hi4)], per_lane=per_lane) as b, cw.Session(dev, lo4, hi4) as s:
            b.run(40000000, 900000000)
            s.run(40000000, 900000000)
            assert b.residues(0) == s.residues()


def test_batch_host_prime_source(dev):
    """Test primes from the host sieve (worker threads) or the device sieve
    build the same batch and give the same results."""
    with cw.Batch(dev, MIX[2:8], host_primes=True) as a, cw.Batch(dev, MIX[2:8]) as b:
        for x, y in zip(a.search(), b.search()):
            _same(x, y)


def test_batch_checkpoint_with_a_finish_in_flight(dev, tmp_path):
    """A checkpoint written while a range's finish is launched but not yet
    collected (its accumulators already iterated in place) keeps that range's
    results, not those values as residues: resumed, every range ends with the
    uninterrupted results."""
    ranges = MIX[2:8]
    with cw.Batch(dev, ranges) as b:
        ref = b.search()
    a, bp = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    with cw.Batch(dev, ranges) as b:
        m2 = b.info[2].maxtarget
        b.run(2, m2 + 1)
        b.finish_launch(2)
        b.checkpoint(a)  # collects the finish first
        assert b.finish_ready(2)
        _same(b.finish_range(2), ref[2])
    with cw.Batch(dev, ranges) as b:
        assert b.resume(a, bp) == (1, 0)
        assert b.covered() == m2 + 1
        b.run(b.covered(), b.maxtarget + 1)
        b.iterate()
        b.finalize()
        for r in range(len(ranges)):
            _same(b.results(r), ref[r])


def test_batch_resume_refused_with_a_finish_in_flight(dev, tmp_path):
    """Reading a checkpoint into a batch whose range has a finish in flight
    is refused (the finish would hand back results of the replaced state);
    the finish then collects the range's own results."""
    ranges = MIX[2:4]
    a, bp = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    with cw.Batch(dev, ranges) as b:
        m0 = b.info[0].maxtarget
        b.run(2, m0 + 1)
        b.checkpoint(a)
        b.finish_launch(0)
        with pytest.raises(cw.ClwError):
            b.resume(a, bp)
        _same(b.finish_range(0), dev.wilson_range(*ranges[0]))


def test_batch_checkpoint_resume(dev, tmp_path):
    """A batch checkpointed part-way (A, then B), rebuilt and resumed from
    the newer valid file, ends with the uninterrupted results; a corrupt B
    falls back to A; a range finished before the checkpoint stays finished."""
    ranges = MIX[2:8]
    with cw.Batch(dev, ranges) as b:
        ref = b.search()
    a, bp = str(tmp_path / "stateA.ckp"), str(tmp_path / "stateB.ckp")
    with cw.Batch(dev, ranges) as b:
        b.run(2, 10_000_000)
        b.checkpoint(a)
        m2 = b.info[2].maxtarget  # (86e6, 86.01e6): target ~4.3e7
        assert m2 > 10_000_000
        b.run(10_000_000, m2 + 1)
        early = [b.finish_range(0), b.finish_range(2)]
        b.checkpoint(bp)
    _same(early[0], ref[0])
    _same(early[1], ref[2])
    with cw.Batch(dev, ranges) as b:
        assert b.resume(a, bp) == (1, 1)
        assert b.covered() == m2 + 1
        b.run(b.covered(), b.maxtarget + 1)
        b.iterate()
        b.finalize()
        for r in range(len(ranges)):
            _same(b.results(r), ref[r])
    raw = bytearray(open(bp, "rb").read())
    raw[-3] ^= 0x40
    open(bp, "wb").write(bytes(raw))
    with cw.Batch(dev, ranges) as b:
        assert b.resume(a, bp) == (1, 0)
        assert b.covered() == 10_000_000
        b.run(b.covered(), b.maxtarget + 1)
        b.iterate()
        b.finalize()
        for r in range(len(ranges)):
            _same(b.results(r), ref[r])
    with cw.Batch(dev, ranges[:3]) as b:  # other ranges: no valid file
        assert b.resume(a, bp)[0] == 0


def test_batch_edge_ranges(dev):
    """Ranges at the edges: only p = 5 (target 1, nothing to sieve), a single
    prime, ranges straddling the 2/3-limb band boundary (7.59e8), a range of
    one type only, all in one batch, each equal to its own run."""
    ranges = [(5, 6), (5, 12), (97, 98), (759250000, 759260000), (759250120, 759250200), (1009, 1014)]
    with cw.Batch(dev, ranges) as b:
        assert [st.limbs for st in b.info][3:5] == [3, 3]
        got = b.search()
    for (lo, hi), r in zip(ranges, got):
        _same(r, dev.wilson_range(lo, hi))


def test_batch_band_boundary_3_4_limbs(dev):
    """Test primes on both sides of the 3/4-limb boundary (2.4879e13): window
    residues equal the session path's."""
    ranges = [(24879108090000, 24879108095803), (24879108095803, 24879108100000)]
    with cw.Batch(dev, ranges) as b:
        assert [st.limbs for st in b.info] == [3, 4]
        b.run(2, 300_000_000)
        for r, (lo, hi) in enumerate(ranges):
            with cw.Session(dev, lo, hi) as s:
                s.run(2, 300_000_000)
                assert b.residues(r) == s.residues()
