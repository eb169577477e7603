"""GPU tests of the request loop (continuous batching): every request's
results are those of a single-range run, whatever else shared the GPU with
it, and the HBM budget holds. Integer work: bit-exact comparisons."""
import sys

import numpy as np
import pytest

import clwilson_amd as cw
from clwilson_amd.serve import Server
from clwilson_amd.workloads import Request

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True, file=sys.stderr)


RANGES = [(1_000_000, 1_005_000), (5_000_000, 5_004_000), (100_000_000, 100_003_000), (20_000, 40_000),
          (750_000_000, 750_002_000), (2_000_000_000, 2_000_002_000), (300_000_000, 300_001_000),
          (1_000_000_000, 1_000_003_000), (7, 100), (60_000_000, 60_005_000), (900_000_000, 900_000_500),
          (3_000_000, 3_000_100)]


def _same(a, b):
    assert np.array_equal(a.p, b.p)
    assert np.array_equal(a.quot, b.quot)
    assert np.array_equal(a.rem, b.rem)
    assert a.checksum == b.checksum
    assert (a.resultcount, a.lastp, a.lastw) == (b.resultcount, b.lastp, b.lastw)
    assert a.results == b.results


@pytest.mark.parametrize("streams", [1, 3])
def test_serving_alone_vs_mix(dev, streams):
    """Twelve requests arriving over 0.3 s under an HBM budget that admits
    about three cohorts at a time: every result equals the request run alone
    (the session path), and the budget is never exceeded."""
    reqs = [Request(i, 0.025 * i, lo, hi) for i, (lo, hi) in enumerate(RANGES)]
    srv = Server(hbm_budget=1 << 40, streams=streams, step_seconds=0.01, admit_interval=0.05, segment_blocks=64)
    budget = srv.budget = srv.used + (30 << 20)  # per-stream scratch and workspace + ~3 cohorts
    try:
        rep = srv.serve(reqs, duration=100.0, until_done=True)
    finally:
        srv.close()
    say(rep)
    assert rep["completed"] == len(reqs)
    assert rep["hbm_peak"] <= budget
    assert rep["p50_latency_s"] is not None and rep["p99_latency_s"] >= rep["p50_latency_s"]
    got = {c.rid: c.result for c in srv.completed}
    for r in reqs:
        say("request", r.rid, r.pmin, r.pmax)
        _same(got[r.rid], dev.wilson_range(r.pmin, r.pmax))


def test_finish_range_early_matches_full_batch(dev):
    """A range finished early (its own iterate + check while the batch goes
    on) equals the same range finished with the whole batch."""
    ranges = [(1_000_000, 1_002_000), (400_000_000, 400_002_000), (50_000, 60_000)]
    with cw.Batch(dev, ranges) as b:
        ref = b.search()
    with cw.Batch(dev, ranges) as b:
        maxt = [st.maxtarget for st in b.info]
        b.run(2, maxt[2] + 1)
        early = b.finish_range(2)
        with pytest.raises(cw.ClwError):
            b.finish_range(1)  # its sieving primes are not all in yet
        b.run(maxt[2] + 1, maxt[0] + 1)
        early0 = b.finish_range(0)
        b.run(maxt[0] + 1, b.maxtarget + 1)
        b.iterate()
        b.finalize()
        _same(early, ref[2])
        _same(early0, ref[0])
        _same(b.results(1), ref[1])
        _same(b.results(2), ref[2])


def test_finish_launch_collect_matches_full_batch(dev):
    """Finishing in two halves: the range's iterate and check queued on the
    stream (finish_launch), a window of the other ranges launched behind it,
    then the results collected once finish_ready says so; they equal the
    whole batch's. A finish of a range whose sieving primes are not all in is
    refused, and so is a query of a range never launched."""
    ranges = [(1_000_000, 1_002_000), (400_000_000, 400_002_000), (50_000, 60_000)]
    with cw.Batch(dev, ranges) as b:
        ref = b.search()
    with cw.Batch(dev, ranges) as b:
        maxt = [st.maxtarget for st in b.info]
        b.run(2, maxt[2] + 1)
        with pytest.raises(cw.ClwError):
            b.finish_ready(2)  # not launched
        b.finish_launch(2)
        b.finish_launch(2)  # a second launch of the same range does nothing
        with pytest.raises(cw.ClwError):
            b.finish_launch(1)  # its sieving primes are not all in yet
        b.launch(maxt[2] + 1, maxt[0] + 1)  # behind the finish on the same stream
        b.wait()
        assert b.finish_ready(2)
        early = b.finish_range(2)
        assert b.finish_ready(2)  # collected: stays ready
        b.finish_launch(0)
        b.run(maxt[0] + 1, b.maxtarget + 1)
        b.iterate()  # skips ranges 0 and 2
        b.finalize()
        early0 = b.finish_range(0)
        _same(early, ref[2])
        _same(early0, ref[0])
        _same(b.results(1), ref[1])
        _same(b.results(2), ref[2])


def test_batches_sharing_a_workspace(dev):
    """Two batches whose windows alternate on one stream in one workspace's
    buffers (record buffer, sieve bitmap, run products on the matrix cores at
    ~2e9) give the results of the same batches with buffers of their own; a
    batch created without window buffers refuses a window until it has a
    workspace."""
    ra = [(2_000_000_000, 2_000_004_000), (1_000_000_000, 1_000_002_000)]
    rb = [(3_000_000_000, 3_000_003_000), (50_000, 60_000)]
    with cw.Batch(dev, ra) as a, cw.Batch(dev, rb) as b:
        ref_a, ref_b = a.search(), b.search()
    ws = cw.Workspace(dev)
    try:
        with cw.Batch(dev, ra, external_workspace=True) as a, cw.Batch(dev, rb, external_workspace=True) as b:
            with pytest.raises(cw.ClwError):
                a.run(2, 1000)
            other = cw.Device(0)  # another context (stream) on the same GPU
            try:
                ws_other = cw.Workspace(other)
                try:
                    with pytest.raises(cw.ClwError):
                        a.set_workspace(ws_other)
                finally:
                    ws_other.close()
                a.set_workspace(ws)
                a.bind(other)  # now on another stream than its workspace
                with pytest.raises(cw.ClwError):
                    a.run(2, 1000)
                a.bind(dev)
            finally:
                other.close()
            a.set_workspace(ws)
            b.set_workspace(ws)
            cuts = [2, 10**8, 4 * 10**8]
            for x in (a, b):
                x.cuts = cuts + [x.maxtarget + 1]
            for k in range(len(cuts)):
                for x in (a, b):
                    say("window", k, x.cuts[k], x.cuts[k + 1])
                    x.run(x.cuts[k], x.cuts[k + 1])
            assert a.kernel_stats()["mfma_int8_macs"] > 0
            for x in (a, b):
                x.iterate()
                x.finalize()
            for r in range(2):
                _same(a.results(r), ref_a[r])
                _same(b.results(r), ref_b[r])
    finally:
        ws.close()


def test_server_snapshot_restore(dev, tmp_path):
    """A server stopped part-way and snapshotted, then restored into a new
    server, completes every request with the single-range results."""
    reqs = [Request(i, 0.0, lo, hi) for i, (lo, hi) in enumerate(RANGES)]
    srv = Server(hbm_budget=1 << 30, streams=2, step_seconds=0.002, segment_blocks=64, admit_wait=0.0)
    try:
        for r in reqs:
            srv.submit(r)
        srv.admit()
        for _ in range(6):
            srv.poll()
        srv.snapshot(str(tmp_path))
        before = {c.rid: c.result for c in srv.completed}
    finally:
        srv.close()
    say("completed before the snapshot:", sorted(before))
    srv2 = Server(hbm_budget=1 << 30, streams=2, step_seconds=0.01, segment_blocks=64, admit_wait=0.0)
    try:
        done = srv2.restore(str(tmp_path))
        assert sorted(done) == sorted(before)
        while not srv2.idle():
            if not srv2.poll():
                pass
        after = {c.rid: c.result for c in srv2.completed}
    finally:
        srv2.close()
    got = {**before, **after}
    assert sorted(got) == [r.rid for r in reqs]
    for r in reqs:
        _same(got[r.rid], dev.wilson_range(r.pmin, r.pmax))
