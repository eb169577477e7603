"""Host logic of the request loop (no GPU): admission against the HBM
budget, magnitude classes and cohort caps, and the latency report."""
import math

import numpy as np

import clwilson_amd.serve as S
from clwilson_amd.workloads import Request


class FakeStats:
    def __init__(self, lo, hi):
        T = max(hi // 2, 2)
        self.typeTarget = [T // 3, T // 2, T]
        self.tpcnt = [1, 1, 1]
        self.maxtarget = T
        self.tpcount = 3


class FakeBatch:
    built = []

    def __init__(self, dev, ranges, **kw):
        self.ranges = ranges
        self.info = [FakeStats(lo, hi) for lo, hi in ranges]
        self.kw = kw
        self.h = 1
        FakeBatch.built.append(self)

    def device_bytes(self):
        return sum(S.Server.estimate_bytes(Request(0, 0.0, lo, hi)) for lo, hi in self.ranges) // 2

    def covered(self):
        return 2

    def close(self):
        self.h = None

    def wait(self):
        pass


class FakeDevice:
    def __init__(self, index=0):
        pass

    def close(self):
        pass


class FakeScratch:
    def __init__(self, dev, cap):
        self.cap = cap

    def close(self):
        pass


class FakeWorkspace:
    def __init__(self, dev, segment_span):
        self.segment_span = segment_span
        self.bytes = 0

    def close(self):
        pass


def make_server(monkeypatch, **kw):
    monkeypatch.setattr(S, "Batch", FakeBatch)
    monkeypatch.setattr(S, "Device", FakeDevice)
    monkeypatch.setattr(S, "Scratch", FakeScratch)
    monkeypatch.setattr(S, "Workspace", FakeWorkspace)
    FakeBatch.built = []
    kw.setdefault("admit_wait", 0.0)
    return S.Server(**kw)


def test_admission_respects_budget_and_classes(monkeypatch):
    srv = make_server(monkeypatch, hbm_budget=3 * S.Server.SCRATCH_SLOTS * 16 + (200 << 20), streams=3)
    reqs = [Request(i, 0.0, 10**9 * (1 + i % 7) * (1000 if i % 2 else 1), 10**9 * (1 + i % 7) *
                    (1000 if i % 2 else 1) + 2_000_000) for i in range(40)]
    for r in reqs:
        srv.submit(r)
    n = srv.admit()
    assert 0 < n < len(reqs)                      # the budget stops admission
    assert srv.used <= srv.budget and srv.reserved == 0
    for b in FakeBatch.built:                     # one magnitude class per cohort
        classes = {S.Server.magnitude_class(Request(0, 0.0, lo, hi)) for lo, hi in b.ranges}
        assert len(classes) == 1
        assert b.kw["external_scratch"] and b.kw["host_primes"] and b.kw["external_workspace"]
    assert [r.rid for r in srv.queue] == [r.rid for r in reqs[n:]]   # arrival order kept


def test_cohort_prime_cap(monkeypatch):
    srv = make_server(monkeypatch, hbm_budget=1 << 40, streams=1)
    wide = [Request(i, 0.0, 10**9 + i * 10**7, 10**9 + i * 10**7 + 9_999_998) for i in range(40)]
    for r in wide:
        srv.submit(r)
    srv.admit()
    for b in FakeBatch.built:
        est = sum(S.Server.estimate_primes(Request(0, 0.0, lo, hi)) for lo, hi in b.ranges)
        assert est <= S.Server.COHORT_PRIMES or len(b.ranges) == 1


def test_report_percentiles(monkeypatch):
    srv = make_server(monkeypatch, streams=1)
    for k, lat in enumerate([5.0, 1.0, 3.0, 2.0, 4.0]):
        srv.completed.append(S.Completion(k, 0.0, 0.0, lat, type("R", (), {"p": np.zeros(2)})()))
    rep = srv.report(5)
    assert rep["completed"] == 5 and rep["p50_latency_s"] == 3.0 and rep["p99_latency_s"] == 5.0
    assert rep["primes_tested"] == 10


def test_magnitude_classes_are_16x_wide():
    c = [S.Server.magnitude_class(Request(0, 0.0, p, p + 2)) for p in (2 * 10**9, 2 * 16 * 10**9, 2 * 256 * 10**9)]
    assert c[0] < c[1] < c[2] and c[1] - c[0] == 1 and c[2] - c[1] == 1
    assert math.isclose(S.Server.estimate_primes(Request(0, 0.0, 10**12, 10**12 + 10**6)),
                        1.1 * 10**6 / math.log(10**12) + 64)


def test_small_class_waits_to_grow(monkeypatch):
    """A magnitude class with few test primes stays queued until its oldest
    request has waited admit_wait seconds; one with enough primes for a
    cohort goes at once; the others keep their arrival order."""
    srv = make_server(monkeypatch, hbm_budget=1 << 40, streams=1, admit_wait=2.0, min_cohort_primes=1 << 20)
    small = [Request(i, 0.0, 10**12 + i * 10**5, 10**12 + i * 10**5 + 50_000) for i in range(3)]
    big = [Request(10 + i, 0.0, 10**9 + i * 10**7, 10**9 + i * 10**7 + 9_999_998) for i in range(3)]
    for r in small + big:
        srv.submit(r)
    clock = [0.5]
    monkeypatch.setattr(srv, "now", lambda: clock[0])
    assert srv.admit() == 3                       # the big class (~1.6 M test primes) only
    assert [r.rid for r in srv.queue] == [0, 1, 2]
    clock[0] = 2.5                                # the small class's oldest has waited 2.5 s
    assert srv.admit() == 3
    assert srv.queue == []
