"""The multi-range workloads are fixed by their seeds: the generated lists
equal the committed fixtures, are disjoint / ordered as specified, and the
prime-counting estimate used for work accounting matches known values."""
import json
import os

from clwilson_amd.workloads import batch_ranges, pi_approx, serving_requests

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "workloads.json")


def test_batch_ranges_fixture():
    g = json.load(open(GOLDEN))
    r = batch_ranges()
    assert [list(x) for x in r] == g["batch_ranges"]
    assert len(r) == 256
    for (a, b), (c, _) in zip(r, r[1:]):
        assert a < b <= c  # disjoint, increasing
    for a, b in r:
        assert 10**9 <= a <= 10**14 + 10**7 and (b - a) % 2 == 0 and 10**5 <= b - a <= 10**7


def test_serving_requests_fixture():
    g = json.load(open(GOLDEN))
    reqs = serving_requests()
    assert [[q.rid, q.arrival, q.pmin, q.pmax] for q in reqs] == g["serving_requests"]
    assert len(reqs) == 2000
    assert all(b.arrival >= a.arrival for a, b in zip(reqs, reqs[1:]))
    assert 30 < reqs[-1].arrival < 50  # mean 50 requests/s
    for q in reqs:
        assert 10**9 <= q.pmin <= 10**14 and (q.pmax - q.pmin) % 2 == 0
        assert 10**4 <= q.pmax - q.pmin <= 10**7


def test_pi_approx_known_values():
    # pi(10^k), k = 6, 9, 12, 13, with the relative error shrinking as x grows
    for x, pi, tol in [(10**6, 78498, 1e-3), (10**9, 50847534, 1e-6), (10**12, 37607912018, 1e-7),
                       (10**13, 346065536839, 1e-7)]:
        assert abs(pi_approx(x) - pi) / pi < tol
